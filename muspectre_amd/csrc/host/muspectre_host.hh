// Host layer of the MI355X build: the reference's C++ interfaces for the
// Newton-CG path (FFTEngineBase, ProjectionBase, Cell/CellBase, materials,
// SolverBase/SolverCG, newton_cg, de_geus), implemented over the C-ABI in
// include/mugpu.h.  Fields live on the device; host arrays appear only at
// the API edges (get_strain, OptimizeResult, the python bindings).
//
// Reference interfaces (paths relative to the reference repository root):
//   src/fft/fft_engine_base.hh:75-145   src/fft/projection_base.hh:95-161
//   src/cell/cell_base.hh:70-526        src/materials/material_base.hh:103-150
//   src/solver/solver_base.hh:49-120    src/solver/solver_cg.hh:51-106
//   src/solver/solvers.hh:54-74         src/solver/solver_common.hh:50-91
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mugpu.h"

namespace muSpectre {

using Real = double;
using Dim_t = int;
using Uint = unsigned int;
using Int = int;

// src/common/common.hh:142-153
enum class Formulation { finite_strain, small_strain };
enum class FFT_PlanFlags { estimate, measure, patient };
enum class StressMeasure { Cauchy, PK1, PK2, Kirchhoff, Biot, Mandel, no_stress_ };
enum class StrainMeasure { Gradient, Infinitesimal, GreenLagrange, Biot, Log, Almansi,
                           RCauchyGreen, LCauchyGreen, no_strain_ };
// precision of the operator stage (no reference counterpart: the reference
// is FP64 only, src/common/common.hh:69)
enum class Precision { fp64, fp32 };

template <Dim_t Dim>
using Ccoord_t = std::array<Dim_t, Dim>;
template <Dim_t Dim>
using Rcoord_t = std::array<Real, Dim>;

// exceptions (src/solver/solver_common.hh:84-91, src/fft/projection_base.hh:50-58)
class SolverError : public std::runtime_error {
  using std::runtime_error::runtime_error;
};
class ConvergenceError : public SolverError {
  using SolverError::SolverError;
};
class ProjectionError : public std::runtime_error {
  using std::runtime_error::runtime_error;
};
class DeviceError : public std::runtime_error {
  using std::runtime_error::runtime_error;
};

// maps a C-ABI return code onto the reference's exception types
void check(int rc);

// dense column-major matrix standing in for Eigen::MatrixXd in the load
// steps (src/solver/solvers.hh:48)
struct Matrix {
  Dim_t rows{0}, cols{0};
  std::vector<Real> data;  // column-major
  Real operator()(Dim_t i, Dim_t j) const { return data[i + rows * j]; }
};
using LoadSteps_t = std::vector<Matrix>;

// a device vector of doubles (the fields and CG vectors)
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(int64_t n);
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  DeviceVector(DeviceVector&& o) noexcept;
  DeviceVector& operator=(DeviceVector&& o) noexcept;
  ~DeviceVector();
  void resize(int64_t n);
  int64_t size() const { return n_; }
  void* data() const { return ptr_; }
  void upload(const Real* h);
  void download(Real* h) const;
  std::vector<Real> to_host() const;
  void set_zero();

 private:
  void* ptr_{nullptr};
  int64_t n_{0};
};

// A field of a result record: a device copy taken when the record is made
// (one HBM-speed copy), downloaded to the host on first access and kept.
class ResultField {
 public:
  ResultField() = default;
  // copies n values from device memory
  void capture(const void* d_src, int64_t n);
  int64_t size() const { return host_valid_ ? (int64_t)host_.size() : (dev_ ? dev_->size() : 0); }
  const std::vector<Real>& host() const;
  void assign(const Real* h, int64_t n);

 private:
  std::shared_ptr<DeviceVector> dev_;
  mutable std::vector<Real> host_;
  mutable bool host_valid_{false};
};

// src/solver/solver_common.hh:50-67
struct OptimizeResult {
  ResultField grad;    // nb_dof values (AoS)
  ResultField stress;  // nb_dof values (AoS)
  bool success{false};
  Int status{0};
  std::string message;
  Uint nb_it{0};
  Uint nb_fev{0};
};

// ---- FFT engine (src/fft/fft_engine_base.hh:75-145, fftw_engine.cc) -------
template <Dim_t Dim>
class FFTEngineBase {
 public:
  using Ccoord = Ccoord_t<Dim>;
  FFTEngineBase(Ccoord resolutions, Dim_t nb_components);
  virtual ~FFTEngineBase() = default;
  FFTEngineBase(const FFTEngineBase&) = delete;
  FFTEngineBase& operator=(const FFTEngineBase&) = delete;
  virtual void initialise(FFT_PlanFlags plan_flags = FFT_PlanFlags::estimate) = 0;
  // r2c of a device field (N x nc AoS) into the engine-owned work space
  virtual DeviceVector& fft(const DeviceVector& field) = 0;
  // c2r of the work space into a device field (the work space is overwritten)
  virtual void ifft(DeviceVector& field) = 0;
  DeviceVector& get_work_space() { return work; }
  Real normalisation() const { return norm_factor; }
  const Ccoord& get_subdomain_resolutions() const { return resolutions; }
  Ccoord get_subdomain_locations() const { return Ccoord{}; }
  const Ccoord& get_domain_resolutions() const { return resolutions; }
  const Ccoord& get_fourier_resolutions() const { return fourier_resolutions; }
  Ccoord get_fourier_locations() const { return Ccoord{}; }
  Dim_t get_nb_components() const { return nb_components; }
  int64_t size() const;
  int64_t fourier_size() const;
  bool is_initialised() const { return initialised; }

 protected:
  Ccoord resolutions;
  Ccoord fourier_resolutions;
  Dim_t nb_components;
  Real norm_factor;
  DeviceVector work;  // Kpts * nc complex, interleaved
  bool initialised{false};
};

// HIP engine: the FFTWEngine replacement (src/fft/fftw_engine.cc:40-158)
template <Dim_t Dim>
class HIPEngine : public FFTEngineBase<Dim> {
 public:
  using Parent = FFTEngineBase<Dim>;
  using typename Parent::Ccoord;
  HIPEngine(Ccoord resolutions, Dim_t nb_components, Precision prec = Precision::fp64);
  ~HIPEngine() override;
  void initialise(FFT_PlanFlags plan_flags = FFT_PlanFlags::estimate) override;
  DeviceVector& fft(const DeviceVector& field) override;
  void ifft(DeviceVector& field) override;
  mugpu_fft handle() const { return plan; }
  // hands the plan over to a projection (src/fft/projection_base.cc:41-53)
  mugpu_fft release();
  Precision get_precision() const { return prec; }

 private:
  mugpu_fft plan{nullptr};
  Precision prec;
};

// ---- projections (src/fft/projection_base.hh:95-161) -----------------------
template <Dim_t DimS, Dim_t DimM = DimS>
class ProjectionBase {
 public:
  using Ccoord = Ccoord_t<DimS>;
  using Rcoord = Rcoord_t<DimS>;
  ProjectionBase(std::unique_ptr<HIPEngine<DimS>> engine, Rcoord lengths, Formulation form);
  virtual ~ProjectionBase();
  virtual void initialise(FFT_PlanFlags flags = FFT_PlanFlags::estimate);
  // in place on a device field (N x nc)
  virtual void apply_projection(DeviceVector& field);
  // host copy of the reference's stored operator, Kpts x (nc*nc) or Kpts x dim
  virtual std::vector<Real> get_operator();
  std::array<Dim_t, 2> get_strain_shape() const { return {DimM, DimM}; }
  Dim_t get_nb_components() const { return DimM * DimM; }
  Formulation get_formulation() const { return form; }
  const Rcoord& get_domain_lengths() const { return lengths; }
  const Ccoord& get_domain_resolutions() const { return resolutions; }
  const Ccoord& get_subdomain_resolutions() const { return resolutions; }
  Ccoord get_subdomain_locations() const { return Ccoord{}; }
  int64_t size() const;
  int64_t fourier_size() const;
  int kind() const { return kind_; }

 protected:
  ProjectionBase(std::unique_ptr<HIPEngine<DimS>> engine, Rcoord lengths, Formulation form,
                 int kind);
  Ccoord resolutions;
  Rcoord lengths;
  Formulation form;
  int kind_;
  mugpu_proj proj{nullptr};
  std::unique_ptr<HIPEngine<DimS>> engine;
  bool initialised{false};
};

template <Dim_t DimS, Dim_t DimM = DimS>
class ProjectionSmallStrain : public ProjectionBase<DimS, DimM> {
 public:
  ProjectionSmallStrain(std::unique_ptr<HIPEngine<DimS>> engine, Rcoord_t<DimS> lengths);
};
template <Dim_t DimS, Dim_t DimM = DimS>
class ProjectionFiniteStrainFast : public ProjectionBase<DimS, DimM> {
 public:
  ProjectionFiniteStrainFast(std::unique_ptr<HIPEngine<DimS>> engine, Rcoord_t<DimS> lengths);
};
template <Dim_t DimS, Dim_t DimM = DimS>
class ProjectionFiniteStrain : public ProjectionBase<DimS, DimM> {
 public:
  ProjectionFiniteStrain(std::unique_ptr<HIPEngine<DimS>> engine, Rcoord_t<DimS> lengths);
};

// ---- cell (src/cell/cell_base.hh:70-526) -------------------------------------
class MaterialBase;

class Cell {
 public:
  virtual ~Cell();
  Cell(const Cell&) = delete;
  Cell& operator=(const Cell&) = delete;
  int64_t get_nb_dof() const { return nb_pixels * nb_components; }
  int64_t size() const { return nb_pixels; }
  Formulation get_formulation() const { return form; }
  Dim_t get_material_dim() const { return dim; }
  Dim_t get_spatial_dim() const { return dim; }
  std::array<Dim_t, 2> get_strain_shape() const { return {dim, dim}; }
  Precision get_precision() const { return prec; }
  bool is_initialised() const;
  void initialise(FFT_PlanFlags flags = FFT_PlanFlags::estimate);
  // device views of F and P (Cell::get_strain_vector / get_stress_vector)
  void* strain_data() const;
  void* stress_data() const;
  std::vector<Real> get_strain();
  std::vector<Real> get_stress() const;
  void set_strain(const std::vector<Real>& F);
  void set_uniform_strain(const Matrix& strain);
  void add_uniform_strain(const Matrix& increment);
  // CellBase::evaluate_stress_tangent (cell_base.cc:129-146)
  void evaluate_stress_tangent();
  void evaluate_stress();
  std::vector<Real> get_tangent() const;
  // CellBase::evaluate_projected_directional_stiffness (cell_base.cc:150-192)
  void evaluate_projected_directional_stiffness(const DeviceVector& dF, DeviceVector& dP);
  // K : dF only
  void tangent_product(const DeviceVector& dF, DeviceVector& dP);
  // CellBase::apply_projection (cell_base.cc:201-208)
  void apply_projection(DeviceVector& v);
  void save_history_variables();
  std::vector<Real> get_globalised_internal_real_array(const std::string& name, Dim_t& nb_comp);
  // host copy of the operator of the cell's projection (as
  // ProjectionBase::get_operator, projection_base.hh:144)
  std::vector<Real> get_projection_operator() const;
  int64_t fourier_size() const;
  MaterialBase& add_material(std::unique_ptr<MaterialBase> mat);
  const std::vector<std::unique_ptr<MaterialBase>>& get_materials() const { return materials; }
  mugpu_cell handle() const { return cell; }
  const std::vector<Dim_t>& get_resolutions() const { return res; }
  const std::vector<Real>& get_lengths() const { return lengths; }

 protected:
  Cell(Dim_t dim, const std::vector<Dim_t>& res, const std::vector<Real>& lengths,
       Formulation form, Precision prec, bool allow_even);
  Dim_t dim;
  Dim_t nb_components;
  int64_t nb_pixels;
  std::vector<Dim_t> res;
  std::vector<Real> lengths;
  Formulation form;
  Precision prec;
  mugpu_cell cell{nullptr};
  std::vector<std::unique_ptr<MaterialBase>> materials;
};

template <Dim_t DimS, Dim_t DimM = DimS>
class CellBase : public Cell {
 public:
  using Ccoord = Ccoord_t<DimS>;
  using Rcoord = Rcoord_t<DimS>;
  CellBase(Ccoord resolutions, Rcoord lengths, Formulation form,
           Precision prec = Precision::fp64, bool allow_even = false);
  Ccoord get_domain_resolutions() const;
  Rcoord get_domain_lengths() const;
  Ccoord get_subdomain_resolutions() const { return get_domain_resolutions(); }
  Ccoord get_subdomain_locations() const { return Ccoord{}; }
  Ccoord get_ccoord(int64_t index) const;
  int64_t get_index(const Ccoord& c) const;
};

// The cell's own projection, as CellBase::get_projection returns it
// (cell_base.hh:323; bind_py_cell.cc:197-198): the projection of the cell's
// formulation on its grid (small strain: ProjectionSmallStrain, finite
// strain: ProjectionFiniteStrainFast), acting through the cell's operator
// (no second plan or work space).
template <Dim_t DimS, Dim_t DimM = DimS>
class CellProjection {
 public:
  explicit CellProjection(CellBase<DimS, DimM>& cell) : cell(cell) {}
  void initialise(FFT_PlanFlags flags = FFT_PlanFlags::estimate) { cell.initialise(flags); }
  void apply_projection(DeviceVector& field) { cell.apply_projection(field); }
  std::vector<Real> get_operator() const { return cell.get_projection_operator(); }
  Formulation get_formulation() const { return cell.get_formulation(); }
  Ccoord_t<DimS> get_domain_resolutions() const { return cell.get_domain_resolutions(); }
  Ccoord_t<DimS> get_subdomain_resolutions() const { return cell.get_domain_resolutions(); }
  Ccoord_t<DimS> get_subdomain_locations() const { return Ccoord_t<DimS>{}; }
  Rcoord_t<DimS> get_domain_lengths() const { return cell.get_domain_lengths(); }
  std::array<Dim_t, 2> get_strain_shape() const { return {DimM, DimM}; }
  Dim_t get_nb_components() const { return DimM * DimM; }
  int64_t size() const { return cell.size(); }
  int64_t fourier_size() const { return cell.fourier_size(); }
  CellBase<DimS, DimM>& get_cell() { return cell; }

 private:
  CellBase<DimS, DimM>& cell;
};

// make_cell (src/cell/cell_factory.hh:93-105)
template <Dim_t DimS, Dim_t DimM = DimS>
std::unique_ptr<CellBase<DimS, DimM>> make_cell(Ccoord_t<DimS> res, Rcoord_t<DimS> lengths,
                                                 Formulation form,
                                                 Precision prec = Precision::fp64,
                                                 bool allow_even = false);

// ---- materials (src/materials/material_base.hh:103-150) ----------------------
class MaterialBase {
 public:
  MaterialBase(Cell& cell, std::string name, int family, Real young, Real poisson, Real tau_y0,
               Real H);
  // MaterialLinearElasticGeneric (material_linear_elastic_generic.cc:40-65):
  // one stiffness in Voigt notation, vsize x vsize
  MaterialBase(Cell& cell, std::string name, const Matrix& C_voigt);
  // a material of no cell, as the reference's MaterialLinearElastic3/4(name)
  // constructors make one (bind_py_material.cc:115,141): it takes pixels
  // (linear indices of a grid it does not know) and exposes its internal
  // fields, and is never evaluated
  MaterialBase(std::string name, int family, Dim_t dim);
  virtual ~MaterialBase() = default;
  const std::string& get_name() const { return name; }
  int get_family() const { return family; }
  size_t size() const { return pixels.size(); }
  // plain add_pixel (material_base.cc:56-58)
  virtual void add_pixel_index(int64_t index);
  // per-pixel parameters (LE2 eigenstrain, LE3/LE4 Young and Poisson)
  void add_pixel_index_with(int64_t index, const std::vector<Real>& params);
  const std::vector<int64_t>& get_pixels() const { return pixels; }
  // hands pixels to the device cell (at Cell::initialise)
  void upload();
  int index() const { return mat; }
  bool has_cell() const { return cell != nullptr; }
  Cell& get_cell();
  Dim_t get_dim() const { return dim; }
  // The material's internal field collection (MaterialMuSpectre::
  // internal_fields, material_muSpectre_base.hh; FieldCollectionBase,
  // field_collection_base.hh:107-122, 282-293), with the reference's field
  // names: LE2 "Eigenstrain", LE3 "local stiffness tensor", LE4 "local first
  // Lame constant" and "local second Lame constant(shear modulus)", HEP1 the
  // state fields "cumulated plastic flow εₚ", "Previous placement gradient
  // Fᵗ", "Previous left Cauchy-Green deformation bₑᵗ", each two fields
  // "<prefix>, sub_field index 0/1".  Names are sorted as the reference's
  // std::map keys are.
  std::vector<std::string> field_names() const;
  std::vector<std::string> statefield_names() const;
  // values of a real field, one row of nb_comp values per pixel of the
  // material (its own pixel order); throws for a name it does not hold
  std::vector<Real> field_values(const std::string& name, Dim_t& nb_comp) const;
  // overwrites a field held on the host (LE2's eigenstrain) before the cell
  // is initialised; the device copy is made at initialisation
  void set_field_values(const std::string& name, const std::vector<Real>& values);
  // the sub-field of a state field that is current (statefield.hh:245-257)
  int current_sub_field(const std::string& prefix) const;

 protected:
  Cell* cell;
  std::string name;
  int family;
  Dim_t dim;
  int mat{-1};
  int nb_params;
  std::vector<int64_t> pixels;
  std::vector<Real> params;
  bool uploaded{false};
};

// ---- solvers (src/solver/solver_base.hh:49-120, solver_cg.hh:51-106) --------
class SolverBase {
 public:
  SolverBase(Cell& cell, Real tol, Uint maxiter, bool verbose = false);
  virtual ~SolverBase() = default;
  virtual void initialise() {}
  // solves A x = rhs on the device; returns the solver-owned solution
  virtual DeviceVector& solve(const DeviceVector& rhs) = 0;
  virtual std::string get_name() const = 0;
  Uint get_counter() const { return counter; }
  void reset_counter() { counter = 0; }
  bool has_converged() const { return converged; }
  Uint get_maxiter() const { return maxiter; }
  Real get_tol() const { return tol; }
  Cell& get_cell() { return cell; }

 protected:
  Cell& cell;
  Real tol;
  Uint maxiter;
  bool verbose;
  Uint counter{0};
  bool converged{false};
};

class SolverCG : public SolverBase {
 public:
  using Parent = SolverBase;
  SolverCG(Cell& cell, Real tol, Uint maxiter, bool verbose = false);
  DeviceVector& solve(const DeviceVector& rhs) override;
  std::string get_name() const override { return "CG"; }
  // FP32 cells: relative tolerance of the inner FP32 CG between two FP64
  // residual refinements (an inner solve stops earlier once the outer
  // tolerance is met with a factor of two to spare)
  void set_inner_tol(Real t) { inner_tol = t; }
  int64_t get_refinements() const { return refinements; }

 protected:
  DeviceVector r_k, p_k, Ap_k, x_k;
  Real inner_tol{1e-4};
  int64_t refinements{0};
};

// ---- Eigen's iterative solvers (src/solver/solver_eigen.hh:58-225) ----------
// The reference binds Eigen 3.3's ConjugateGradient, BiCGSTAB, GMRES, DGMRES
// and MINRES, each with the identity preconditioner, to the cell through a
// sparse-matrix adaptor whose product is the projected directional
// stiffness (cell_base.hh:534-591).  Here the same algorithms run on the
// device operator and device vectors; the host keeps the scalar recurrences
// (and, for GMRES, the small Hessenberg system).  solve() follows
// SolverEigen::solve (solver_eigen.cc:61-86): zero initial guess, counter +=
// iterations, ConvergenceError when the relative residual stays above tol.
class SolverEigenBase : public SolverBase {
 public:
  SolverEigenBase(Cell& cell, Real tol, Uint maxiter, bool verbose = false);
  DeviceVector& solve(const DeviceVector& rhs) override;

 protected:
  // one solve from x = 0; iters and the relative residual error out;
  // false on a numerical breakdown
  virtual bool run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) = 0;
  void apply(const DeviceVector& in, DeviceVector& out);  // A = G K (projected)
  Real dot(const DeviceVector& a, const DeviceVector& b) const;
  // y = a x + b y
  void axpby(Real a, const DeviceVector& x, Real b, DeviceVector& y) const;
  // y = a x
  void scale_copy(Real a, const DeviceVector& x, DeviceVector& y) const;
  DeviceVector x_k;
};

class SolverCGEigen : public SolverEigenBase {
 public:
  using SolverEigenBase::SolverEigenBase;
  std::string get_name() const override { return "CG"; }

 protected:
  bool run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) override;
};

class SolverBiCGSTABEigen : public SolverEigenBase {
 public:
  using SolverEigenBase::SolverEigenBase;
  std::string get_name() const override { return "BiCGSTAB"; }

 protected:
  bool run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) override;
};

class SolverGMRESEigen : public SolverEigenBase {
 public:
  using SolverEigenBase::SolverEigenBase;
  std::string get_name() const override { return "GMRES"; }

 protected:
  bool run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) override;
  Uint restart{30};  // Eigen's GMRES and DGMRES default
};

// Eigen's DGMRES with its default settings, which deflate no eigenvalues:
// the restarted GMRES(30) of SolverGMRESEigen (solvers_krylov.cc states the
// correspondence and the one deliberate difference, the start x0 = 0)
class SolverDGMRESEigen : public SolverGMRESEigen {
 public:
  using SolverGMRESEigen::SolverGMRESEigen;
  std::string get_name() const override { return "DGMRES"; }
};

class SolverMINRESEigen : public SolverEigenBase {
 public:
  using SolverEigenBase::SolverEigenBase;
  std::string get_name() const override { return "MINRES"; }

 protected:
  bool run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) override;
};

// src/solver/solvers.hh:54-74
// newton_maxiter: optional bound on the Newton iterations per load step
// (no reference counterpart; 0 keeps the reference's bound, the solver's
// maxiter, src/solver/solvers.cc:165-168)
std::vector<OptimizeResult> newton_cg(Cell& cell, const LoadSteps_t& load_steps,
                                      SolverBase& solver, Real newton_tol, Real equil_tol,
                                      Dim_t verbose = 0, Uint newton_maxiter = 0);
std::vector<OptimizeResult> de_geus(Cell& cell, const LoadSteps_t& load_steps,
                                    SolverBase& solver, Real newton_tol, Real equil_tol,
                                    Dim_t verbose = 0, Uint newton_maxiter = 0);

// src/solver/solver_common.cc:39-43
bool check_symmetry(const Matrix& eps, Real rel_tol = 1e-8);

}  // namespace muSpectre
