// Implementation of the host layer over the C-ABI (include/mugpu.h).
#include "muspectre_host.hh"

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace muSpectre {

void check(int rc) {
  if (rc == MUGPU_OK) return;
  const std::string msg = mugpu_last_error();
  switch (rc) {
    case MUGPU_ECONVERGENCE: throw ConvergenceError(msg);
    case MUGPU_ESOLVER: throw SolverError(msg);
    case MUGPU_EPROJECTION: throw ProjectionError(msg);
    case MUGPU_ERUNTIME:
    case MUGPU_EARG: throw std::runtime_error(msg);
    default: throw DeviceError(msg);
  }
}

// ---- DeviceVector --------------------------------------------------------------
DeviceVector::DeviceVector(int64_t n) { resize(n); }
DeviceVector::DeviceVector(DeviceVector&& o) noexcept : ptr_(o.ptr_), n_(o.n_) {
  o.ptr_ = nullptr;
  o.n_ = 0;
}
DeviceVector& DeviceVector::operator=(DeviceVector&& o) noexcept {
  if (this != &o) {
    if (ptr_) mugpu_free(ptr_);
    ptr_ = o.ptr_;
    n_ = o.n_;
    o.ptr_ = nullptr;
    o.n_ = 0;
  }
  return *this;
}
DeviceVector::~DeviceVector() {
  if (ptr_) mugpu_free(ptr_);
}
void DeviceVector::resize(int64_t n) {
  if (n == n_ && ptr_) return;
  if (ptr_) mugpu_free(ptr_);
  ptr_ = nullptr;
  n_ = n;
  check(mugpu_alloc(sizeof(Real) * (size_t)(n > 0 ? n : 1), &ptr_));
}
void DeviceVector::upload(const Real* h) { check(mugpu_upload(ptr_, h, sizeof(Real) * n_)); }
void DeviceVector::download(Real* h) const { check(mugpu_download(h, ptr_, sizeof(Real) * n_)); }
std::vector<Real> DeviceVector::to_host() const {
  std::vector<Real> v((size_t)n_);
  download(v.data());
  return v;
}
void DeviceVector::set_zero() { check(mugpu_memset_zero(ptr_, sizeof(Real) * n_)); }

// ---- ResultField ---------------------------------------------------------------
void ResultField::capture(const void* d_src, int64_t n) {
  dev_ = std::make_shared<DeviceVector>(n);
  check(mugpu_scale_copy(1.0, d_src, dev_->data(), n));
  host_.clear();
  host_valid_ = false;
}
const std::vector<Real>& ResultField::host() const {
  if (!host_valid_) {
    host_.assign((size_t)(dev_ ? dev_->size() : 0), 0.0);
    if (dev_) dev_->download(host_.data());
    host_valid_ = true;
  }
  return host_;
}
void ResultField::assign(const Real* h, int64_t n) {
  dev_.reset();
  host_.assign(h, h + n);
  host_valid_ = true;
}

// ---- FFT engine ------------------------------------------------------------------
template <Dim_t Dim>
FFTEngineBase<Dim>::FFTEngineBase(Ccoord res, Dim_t nbc)
    : resolutions(res), fourier_resolutions(res), nb_components(nbc) {
  fourier_resolutions[Dim - 1] = res[Dim - 1] / 2 + 1;
  norm_factor = 1.0 / (Real)size();
}
template <Dim_t Dim>
int64_t FFTEngineBase<Dim>::size() const {
  int64_t n = 1;
  for (auto r : resolutions) n *= r;
  return n;
}
template <Dim_t Dim>
int64_t FFTEngineBase<Dim>::fourier_size() const {
  int64_t n = 1;
  for (auto r : fourier_resolutions) n *= r;
  return n;
}

template <Dim_t Dim>
HIPEngine<Dim>::HIPEngine(Ccoord res, Dim_t nbc, Precision p) : Parent(res, nbc), prec(p) {}

template <Dim_t Dim>
HIPEngine<Dim>::~HIPEngine() {
  if (plan) mugpu_fft_destroy(plan);
}

// src/fft/fftw_engine.cc:52-117 (no planning flags are needed on the device)
template <Dim_t Dim>
void HIPEngine<Dim>::initialise(FFT_PlanFlags) {
  if (this->initialised) throw std::runtime_error("double initialisation, will leak memory");
  check(mugpu_fft_create(Dim, this->resolutions.data(), this->nb_components,
                         prec == Precision::fp64 ? MUGPU_FP64 : MUGPU_FP32, &plan));
  this->work.resize(this->fourier_size() * this->nb_components * 2);
  this->initialised = true;
}

template <Dim_t Dim>
DeviceVector& HIPEngine<Dim>::fft(const DeviceVector& field) {
  if (!plan) throw std::runtime_error("fft plan not initialised");
  if (field.size() < this->size() * this->nb_components) throw std::runtime_error("size mismatch");
  check(mugpu_fft_r2c(plan, field.data(), this->work.data()));
  return this->work;
}

template <Dim_t Dim>
void HIPEngine<Dim>::ifft(DeviceVector& field) {
  if (!plan) throw std::runtime_error("ifft plan not initialised");
  if (field.size() < this->size() * this->nb_components) throw std::runtime_error("size mismatch");
  check(mugpu_fft_c2r(plan, this->work.data(), field.data()));
}

template <Dim_t Dim>
mugpu_fft HIPEngine<Dim>::release() {
  mugpu_fft p = plan;
  plan = nullptr;
  return p;
}

template class FFTEngineBase<2>;
template class FFTEngineBase<3>;
template class HIPEngine<2>;
template class HIPEngine<3>;

// ---- projections -------------------------------------------------------------------
template <Dim_t DimS, Dim_t DimM>
ProjectionBase<DimS, DimM>::ProjectionBase(std::unique_ptr<HIPEngine<DimS>> eng, Rcoord L,
                                           Formulation f, int kind)
    : resolutions(eng->get_domain_resolutions()), lengths(L), form(f), kind_(kind),
      engine(std::move(eng)) {
  // src/fft/projection_base.cc:41-53 and the odd-size check of the
  // projection constructors (projection_small_strain.cc:46-51)
  if (engine->get_nb_components() != DimM * DimM)
    throw ProjectionError("Incompatible number of components in FFT engine");
  for (auto r : resolutions)
    if (r % 2 == 0)
      throw ProjectionError("Only an odd number of gridpoints in each direction is supported");
}

template <Dim_t DimS, Dim_t DimM>
ProjectionBase<DimS, DimM>::ProjectionBase(std::unique_ptr<HIPEngine<DimS>> eng, Rcoord L,
                                           Formulation f)
    : ProjectionBase(std::move(eng), L, f,
                     f == Formulation::small_strain ? MUGPU_PROJ_SMALL_STRAIN
                                                    : MUGPU_PROJ_FINITE_STRAIN_FAST) {}

template <Dim_t DimS, Dim_t DimM>
ProjectionBase<DimS, DimM>::~ProjectionBase() {
  if (proj) mugpu_projection_destroy(proj);
}

template <Dim_t DimS, Dim_t DimM>
void ProjectionBase<DimS, DimM>::initialise(FFT_PlanFlags flags) {
  if (initialised) return;
  if (!engine->is_initialised()) engine->initialise(flags);
  mugpu_fft plan = engine->release();
  const int rc = mugpu_projection_create(plan, kind_, lengths.data(), 0, &proj);
  if (rc) mugpu_fft_destroy(plan);
  check(rc);
  initialised = true;
}

template <Dim_t DimS, Dim_t DimM>
void ProjectionBase<DimS, DimM>::apply_projection(DeviceVector& field) {
  if (!initialised) initialise();
  if (field.size() < size() * DimM * DimM) throw std::runtime_error("size mismatch");
  check(mugpu_projection_apply(proj, field.data()));
}

template <Dim_t DimS, Dim_t DimM>
std::vector<Real> ProjectionBase<DimS, DimM>::get_operator() {
  if (!initialised) initialise();
  const int nop = kind_ == MUGPU_PROJ_FINITE_STRAIN_FAST ? DimM : DimM * DimM * DimM * DimM;
  std::vector<Real> op((size_t)(fourier_size() * nop));
  check(mugpu_projection_get_operator(proj, op.data()));
  return op;
}

template <Dim_t DimS, Dim_t DimM>
int64_t ProjectionBase<DimS, DimM>::size() const {
  int64_t n = 1;
  for (auto r : resolutions) n *= r;
  return n;
}
template <Dim_t DimS, Dim_t DimM>
int64_t ProjectionBase<DimS, DimM>::fourier_size() const {
  int64_t n = 1;
  for (int d = 0; d < DimS; ++d) n *= (d == DimS - 1 ? resolutions[d] / 2 + 1 : resolutions[d]);
  return n;
}

template <Dim_t DimS, Dim_t DimM>
ProjectionSmallStrain<DimS, DimM>::ProjectionSmallStrain(std::unique_ptr<HIPEngine<DimS>> e,
                                                         Rcoord_t<DimS> L)
    : ProjectionBase<DimS, DimM>(std::move(e), L, Formulation::small_strain,
                                 MUGPU_PROJ_SMALL_STRAIN) {}
template <Dim_t DimS, Dim_t DimM>
ProjectionFiniteStrainFast<DimS, DimM>::ProjectionFiniteStrainFast(
    std::unique_ptr<HIPEngine<DimS>> e, Rcoord_t<DimS> L)
    : ProjectionBase<DimS, DimM>(std::move(e), L, Formulation::finite_strain,
                                 MUGPU_PROJ_FINITE_STRAIN_FAST) {}
template <Dim_t DimS, Dim_t DimM>
ProjectionFiniteStrain<DimS, DimM>::ProjectionFiniteStrain(std::unique_ptr<HIPEngine<DimS>> e,
                                                           Rcoord_t<DimS> L)
    : ProjectionBase<DimS, DimM>(std::move(e), L, Formulation::finite_strain,
                                 MUGPU_PROJ_FINITE_STRAIN) {}

template class ProjectionBase<2, 2>;
template class ProjectionBase<3, 3>;
template class ProjectionSmallStrain<2, 2>;
template class ProjectionSmallStrain<3, 3>;
template class ProjectionFiniteStrainFast<2, 2>;
template class ProjectionFiniteStrainFast<3, 3>;
template class ProjectionFiniteStrain<2, 2>;
template class ProjectionFiniteStrain<3, 3>;

// ---- cell ------------------------------------------------------------------------
Cell::Cell(Dim_t d, const std::vector<Dim_t>& r, const std::vector<Real>& L, Formulation f,
           Precision p, bool allow_even)
    : dim(d), nb_components(d * d), res(r), lengths(L), form(f), prec(p) {
  nb_pixels = 1;
  for (auto x : r) nb_pixels *= x;
  check(mugpu_cell_create(d, r.data(), L.data(),
                          f == Formulation::small_strain ? MUGPU_SMALL_STRAIN : MUGPU_FINITE_STRAIN,
                          p == Precision::fp64 ? MUGPU_FP64 : MUGPU_FP32, allow_even ? 1 : 0,
                          &cell));
}

Cell::~Cell() {
  materials.clear();
  if (cell) mugpu_cell_destroy(cell);
}

bool Cell::is_initialised() const { return mugpu_cell_is_initialised(cell) != 0; }

// CellBase::initialise (src/cell/cell_base.cc:445-454)
void Cell::initialise(FFT_PlanFlags) {
  if (is_initialised()) return;
  for (auto& m : materials) m->upload();
  check(mugpu_cell_initialise(cell));
}

void* Cell::strain_data() const {
  void* p = nullptr;
  check(mugpu_cell_strain(cell, &p));
  return p;
}
void* Cell::stress_data() const {
  void* p = nullptr;
  check(mugpu_cell_stress(cell, &p));
  return p;
}

// CellBase::get_strain initialises the cell first (src/cell/cell_base.cc:306-313)
std::vector<Real> Cell::get_strain() {
  initialise();
  std::vector<Real> h((size_t)get_nb_dof());
  check(mugpu_download(h.data(), strain_data(), h.size() * sizeof(Real)));
  return h;
}
std::vector<Real> Cell::get_stress() const {
  std::vector<Real> h((size_t)get_nb_dof());
  check(mugpu_download(h.data(), stress_data(), h.size() * sizeof(Real)));
  return h;
}
void Cell::set_strain(const std::vector<Real>& F) {
  if ((int64_t)F.size() != get_nb_dof()) throw std::runtime_error("Size mismatch");
  check(mugpu_upload(strain_data(), F.data(), F.size() * sizeof(Real)));
}
void Cell::set_uniform_strain(const Matrix& s) { check(mugpu_cell_set_uniform_strain(cell, s.data.data())); }
void Cell::add_uniform_strain(const Matrix& s) { check(mugpu_cell_add_uniform_strain(cell, s.data.data())); }

void Cell::evaluate_stress_tangent() {
  initialise();
  check(mugpu_cell_evaluate_stress_tangent(cell));
}
void Cell::evaluate_stress() {
  initialise();
  check(mugpu_cell_evaluate_stress(cell));
}
std::vector<Real> Cell::get_tangent() const {
  std::vector<Real> K((size_t)(nb_pixels * nb_components * nb_components));
  check(mugpu_cell_get_tangent(cell, K.data()));
  return K;
}

void Cell::evaluate_projected_directional_stiffness(const DeviceVector& dF, DeviceVector& dP) {
  if (dF.size() != get_nb_dof()) {
    std::stringstream err;
    err << "input should be of size ndof = " << get_nb_dof() << " but I got " << dF.size();
    throw std::runtime_error(err.str());
  }
  dP.resize(get_nb_dof());
  check(mugpu_cell_directional_stiffness(cell, dF.data(), dP.data()));
}

void Cell::tangent_product(const DeviceVector& dF, DeviceVector& dP) {
  dP.resize(get_nb_dof());
  check(mugpu_cell_tangent_product(cell, dF.data(), dP.data()));
}

void Cell::apply_projection(DeviceVector& v) {
  initialise();
  check(mugpu_cell_apply_projection(cell, v.data()));
}

void Cell::save_history_variables() { check(mugpu_cell_save_history(cell)); }

std::vector<Real> Cell::get_globalised_internal_real_array(const std::string& name, Dim_t& nb) {
  initialise();
  std::vector<Real> out((size_t)(nb_pixels * nb_components));
  int n = 0;
  check(mugpu_cell_get_internal(cell, name.c_str(), out.data(), &n));
  out.resize((size_t)(nb_pixels * n));
  nb = n;
  return out;
}

std::vector<Real> Cell::get_projection_operator() const {
  const int nop = form == Formulation::finite_strain ? dim : nb_components * nb_components;
  std::vector<Real> op((size_t)(fourier_size() * nop));
  check(mugpu_cell_projection_operator(cell, op.data()));
  return op;
}

int64_t Cell::fourier_size() const {
  int64_t n = 1;
  for (Dim_t d = 0; d < dim; ++d) n *= d == dim - 1 ? res[d] / 2 + 1 : res[d];
  return n;
}

MaterialBase& Cell::add_material(std::unique_ptr<MaterialBase> mat) {
  materials.push_back(std::move(mat));
  return *materials.back();
}

template <Dim_t DimS, Dim_t DimM>
CellBase<DimS, DimM>::CellBase(Ccoord r, Rcoord L, Formulation f, Precision p, bool allow_even)
    : Cell(DimS, std::vector<Dim_t>(r.begin(), r.end()), std::vector<Real>(L.begin(), L.end()), f,
           p, allow_even) {}

template <Dim_t DimS, Dim_t DimM>
typename CellBase<DimS, DimM>::Ccoord CellBase<DimS, DimM>::get_domain_resolutions() const {
  Ccoord c;
  for (int d = 0; d < DimS; ++d) c[d] = res[d];
  return c;
}
template <Dim_t DimS, Dim_t DimM>
typename CellBase<DimS, DimM>::Rcoord CellBase<DimS, DimM>::get_domain_lengths() const {
  Rcoord c;
  for (int d = 0; d < DimS; ++d) c[d] = lengths[d];
  return c;
}
// src/common/ccoord_operations.hh:151-192: row-major, last axis fastest
template <Dim_t DimS, Dim_t DimM>
typename CellBase<DimS, DimM>::Ccoord CellBase<DimS, DimM>::get_ccoord(int64_t index) const {
  Ccoord c;
  for (int d = DimS - 1; d >= 0; --d) {
    c[d] = (Dim_t)(index % res[d]);
    index /= res[d];
  }
  return c;
}
template <Dim_t DimS, Dim_t DimM>
int64_t CellBase<DimS, DimM>::get_index(const Ccoord& c) const {
  int64_t i = 0;
  for (int d = 0; d < DimS; ++d) i = i * res[d] + c[d];
  return i;
}

template class CellBase<2, 2>;
template class CellBase<3, 3>;

template <Dim_t DimS, Dim_t DimM>
std::unique_ptr<CellBase<DimS, DimM>> make_cell(Ccoord_t<DimS> res, Rcoord_t<DimS> lengths,
                                                 Formulation form, Precision prec,
                                                 bool allow_even) {
  return std::unique_ptr<CellBase<DimS, DimM>>(
      new CellBase<DimS, DimM>(res, lengths, form, prec, allow_even));
}
template std::unique_ptr<CellBase<2, 2>> make_cell<2, 2>(Ccoord_t<2>, Rcoord_t<2>, Formulation,
                                                         Precision, bool);
template std::unique_ptr<CellBase<3, 3>> make_cell<3, 3>(Ccoord_t<3>, Rcoord_t<3>, Formulation,
                                                         Precision, bool);

// ---- materials -------------------------------------------------------------------
MaterialBase::MaterialBase(Cell& c, std::string n, int fam, Real E, Real nu, Real tau_y0, Real H)
    : cell(&c), name(std::move(n)), family(fam), dim(c.get_material_dim()) {
  check(mugpu_cell_add_material(c.handle(), fam, E, nu, tau_y0, H, &mat));
  const int nc = c.get_material_dim() * c.get_material_dim();
  nb_params = fam == MUGPU_LE2 ? nc : (fam == MUGPU_LE3 || fam == MUGPU_LE4) ? 2 : 0;
}

MaterialBase::MaterialBase(std::string n, int fam, Dim_t d)
    : cell(nullptr), name(std::move(n)), family(fam), dim(d) {
  nb_params = fam == MUGPU_LE2 ? d * d : (fam == MUGPU_LE3 || fam == MUGPU_LE4) ? 2 : 0;
}

Cell& MaterialBase::get_cell() {
  if (!cell) throw std::runtime_error("the material '" + name + "' belongs to no cell");
  return *cell;
}

MaterialBase::MaterialBase(Cell& c, std::string n, const Matrix& C_voigt)
    : cell(&c), name(std::move(n)), family(MUGPU_LEG), dim(c.get_material_dim()), nb_params(0) {
  const Dim_t vs = c.get_material_dim() == 2 ? 3 : 6;
  if (C_voigt.rows != vs || C_voigt.cols != vs) {
    // the reference builds this message (material_linear_elastic_generic.cc:
    // 51-58) but goes on with the wrong-sized matrix; here it is raised
    std::stringstream err;
    err << "The stiffness tensor should be input as a " << vs << " × " << vs
        << " Matrix in Voigt notation. You supplied a " << C_voigt.rows << " × " << C_voigt.cols
        << " matrix";
    throw std::runtime_error(err.str());
  }
  std::vector<Real> rm((size_t)vs * vs);
  for (Dim_t i = 0; i < vs; ++i)
    for (Dim_t j = 0; j < vs; ++j) rm[(size_t)i * vs + j] = C_voigt(i, j);
  check(mugpu_cell_add_material_generic(c.handle(), rm.data(), &mat));
}

void MaterialBase::add_pixel_index(int64_t index) {
  if (nb_params) {
    throw std::runtime_error(family == MUGPU_LE2
                                 ? "this material needs pixels with and eigenstrain"
                                 : "this material needs pixels with Youngs modulus and Poisson ratio.");
  }
  if (uploaded) throw std::runtime_error("pixels must be added before the cell is initialised");
  pixels.push_back(index);
}

void MaterialBase::add_pixel_index_with(int64_t index, const std::vector<Real>& p) {
  if ((int)p.size() != nb_params) throw std::runtime_error("wrong number of per-pixel parameters");
  if (uploaded) throw std::runtime_error("pixels must be added before the cell is initialised");
  pixels.push_back(index);
  params.insert(params.end(), p.begin(), p.end());
}

void MaterialBase::upload() {
  if (uploaded) return;
  check(mugpu_material_add_pixels(get_cell().handle(), mat, (int64_t)pixels.size(), pixels.data(),
                                  nb_params ? params.data() : nullptr));
  uploaded = true;
}

namespace {
// the reference's internal field names (material_linear_elastic2.cc:44,
// material_linear_elastic3.cc:45, material_linear_elastic4.cc:47-50,
// material_hyper_elasto_plastic1.cc:46-53)
const char* const kEigen = "Eigenstrain";
const char* const kLocalC = "local stiffness tensor";
const char* const kLambda = "local first Lame constant";
const char* const kMu = "local second Lame constant(shear modulus)";
const char* const kStatePrefixes[3] = {"cumulated plastic flow \xce\xb5\xe2\x82\x9a",
                                       "Previous placement gradient F\xe1\xb5\x97",
                                       "Previous left Cauchy-Green deformation b\xe2\x82\x91\xe1\xb5\x97"};
// names of a state field's sub-fields (statefield.hh:181-190)
std::string sub_field_name(const std::string& prefix, int i) {
  std::stringstream n;
  n << prefix << ", sub_field index " << i;
  return n.str();
}
bool is_hep(int family) { return family == MUGPU_HEP1 || family == MUGPU_HEP1_CT; }
// Hooke::compute_lambda / compute_mu (materials_toolbox.hh:488-525,603-619)
Real lame_lambda(Real E, Real nu) { return E * nu / ((1 + nu) * (1 - 2 * nu)); }
Real lame_mu(Real E, Real nu) { return E / (2 * (1 + nu)); }
}  // namespace

std::vector<std::string> MaterialBase::field_names() const {
  std::vector<std::string> out;
  if (family == MUGPU_LE2) out.push_back(kEigen);
  if (family == MUGPU_LE3) out.push_back(kLocalC);
  if (family == MUGPU_LE4) {
    out.push_back(kLambda);
    out.push_back(kMu);
  }
  if (is_hep(family))
    for (const char* p : kStatePrefixes)
      for (int i = 0; i < 2; ++i) out.push_back(sub_field_name(p, i));
  std::sort(out.begin(), out.end());
  return out;
}

std::vector<std::string> MaterialBase::statefield_names() const {
  std::vector<std::string> out;
  if (is_hep(family))
    for (const char* p : kStatePrefixes) out.push_back(p);
  std::sort(out.begin(), out.end());
  return out;
}

std::vector<Real> MaterialBase::field_values(const std::string& fname, Dim_t& nb_comp) const {
  const size_t np = pixels.size();
  const Dim_t nc = dim * dim;
  std::vector<Real> out;
  if (family == MUGPU_LE2 && fname == kEigen) {
    nb_comp = nc;
    return params;
  }
  if (family == MUGPU_LE3 && fname == kLocalC) {
    // Hooke::compute_C (materials_toolbox.hh:639-642): lambda I (x) I + 2 mu I4S,
    // entry (i, j, k, l) at i + d j + d^2 k + d^3 l
    nb_comp = nc * nc;
    out.resize(np * (size_t)(nc * nc));
    for (size_t q = 0; q < np; ++q) {
      const Real la = lame_lambda(params[2 * q], params[2 * q + 1]);
      const Real mu = lame_mu(params[2 * q], params[2 * q + 1]);
      Real* C = out.data() + q * (size_t)(nc * nc);
      for (Dim_t l = 0; l < dim; ++l)
        for (Dim_t k = 0; k < dim; ++k)
          for (Dim_t j = 0; j < dim; ++j)
            for (Dim_t i = 0; i < dim; ++i) {
              const Real itrac = (i == j && k == l) ? 1.0 : 0.0;
              const Real isym = 0.5 * (((i == k && j == l) ? 1.0 : 0.0) + ((i == l && j == k) ? 1.0 : 0.0));
              C[(i + dim * j) + nc * (k + dim * l)] = la * itrac + 2 * mu * isym;
            }
    }
    return out;
  }
  if (family == MUGPU_LE4 && (fname == kLambda || fname == kMu)) {
    nb_comp = 1;
    out.resize(np);
    for (size_t q = 0; q < np; ++q)
      out[q] = fname == kLambda ? lame_lambda(params[2 * q], params[2 * q + 1])
                                : lame_mu(params[2 * q], params[2 * q + 1]);
    return out;
  }
  if (is_hep(family)) {
    for (const char* p : kStatePrefixes)
      for (int i = 0; i < 2; ++i)
        if (fname == sub_field_name(p, i)) {
          if (!cell || !cell->is_initialised())
            throw std::runtime_error("the field '" + fname +
                                     "' holds values once the cell is initialised");
          int k = 0, cur = 0;
          check(mugpu_material_get_state(cell->handle(), mat, p, i, nullptr, &k, &cur));
          out.resize(np * (size_t)k);
          check(mugpu_material_get_state(cell->handle(), mat, p, i, out.data(), &k, &cur));
          nb_comp = k;
          return out;
        }
  }
  throw std::runtime_error("The field named '" + fname + "' does not exist in the material '" +
                           name + "'");
}

void MaterialBase::set_field_values(const std::string& fname, const std::vector<Real>& values) {
  if (!(family == MUGPU_LE2 && fname == kEigen))
    throw std::runtime_error("the field '" + fname + "' of this build's material '" + name +
                             "' is derived from its parameters and cannot be assigned");
  if (uploaded)
    throw std::runtime_error("the field '" + fname +
                             "' lives on the device once the cell is initialised");
  if (values.size() != params.size())
    throw std::runtime_error("wrong number of values for the field '" + fname + "'");
  params = values;
}

int MaterialBase::current_sub_field(const std::string& prefix) const {
  if (is_hep(family))
    for (const char* p : kStatePrefixes)
      if (prefix == p) {
        if (!cell || !cell->is_initialised())
          throw std::runtime_error("the state field '" + prefix +
                                   "' holds values once the cell is initialised");
        int k = 0, cur = 0;
        check(mugpu_material_get_state(cell->handle(), mat, p, 0, nullptr, &k, &cur));
        return cur;
      }
  throw std::runtime_error("The state field named '" + prefix +
                           "' does not exist in the material '" + name + "'");
}

// ---- solvers -----------------------------------------------------------------------
SolverBase::SolverBase(Cell& c, Real t, Uint m, bool v) : cell(c), tol(t), maxiter(m), verbose(v) {}

SolverCG::SolverCG(Cell& c, Real t, Uint m, bool v)
    : SolverBase(c, t, m, v), r_k(c.get_nb_dof()), p_k(c.get_nb_dof()), Ap_k(c.get_nb_dof()),
      x_k(c.get_nb_dof()) {}

// src/solver/solver_cg.cc:54-110, device-resident
DeviceVector& SolverCG::solve(const DeviceVector& rhs) {
  const mugpu_cell h = cell.handle();
  converged = false;
  if (cell.get_precision() == Precision::fp32) {
    int64_t its = 0;
    int refs = 0;
    double rel = 0;
    const int rc = mugpu_cg_solve_mixed(h, rhs.data(), x_k.data(), tol, (int)maxiter, inner_tol,
                                        &its, &refs, &rel);
    counter += (Uint)its;
    refinements += refs;
    if (verbose)
      std::cout << "  mixed-precision CG: " << its << " FP32 iterations, " << refs
                << " FP64 residuals, |r|/|b| = " << rel << ", cg_tol = " << tol << std::endl;
    if (rc == MUGPU_ECONVERGENCE) {
      std::stringstream err;
      err << "Conjugate gradient has not converged. After " << counter << " steps, the solver "
          << " FAILED with  |r|/|b| = " << std::setw(15) << rel << ", cg_tol = " << tol
          << std::endl;
      throw ConvergenceError(err.str());
    }
    check(rc);
    converged = true;
    return x_k;
  }
  double rdr = 0, rhs_norm2 = 0;
  check(mugpu_cg_start(h, rhs.data(), x_k.data(), r_k.data(), p_k.data(), Ap_k.data(), &rdr,
                       &rhs_norm2));
  const Real tol2 = tol * tol * rhs_norm2;
  size_t count_width = 0;
  if (verbose) count_width = size_t(std::log10(maxiter)) + 1;
  // for (i = 0; i < maxiter && (rdr > tol2 || i == 0); ++i, ++counter),
  // with the iterations queued on the device kChunk at a time and the loop
  // test applied there after each (one at a time when verbose, to print
  // every step)
  constexpr Uint kChunk = 8;
  for (Uint i = 0; i < maxiter && (rdr > tol2 || i == 0);) {
    const Uint k = verbose ? 1 : std::min<Uint>(kChunk, maxiter - i);
    double new_rdr = 0;
    int done = 0;
    check(mugpu_cg_steps(h, i > 0 ? 1 : 0, x_k.data(), r_k.data(), p_k.data(), Ap_k.data(), tol2,
                         (int)k, &done, &new_rdr));
    rdr = new_rdr;
    i += (Uint)done;
    counter += (Uint)done;
    if (verbose)
      std::cout << "  at CG step " << std::setw(count_width) << i - 1
                << ": |r|/|b| = " << std::setw(15) << std::sqrt(rdr / rhs_norm2)
                << ", cg_tol = " << tol << std::endl;
  }
  // the x update of the last iteration (left pending by mugpu_cg_steps)
  check(mugpu_cg_finish(h));
  if (rdr < tol2) {
    converged = true;
  } else {
    std::stringstream err;
    err << " After " << counter << " steps, the solver "
        << " FAILED with  |r|/|b| = " << std::setw(15) << std::sqrt(rdr / rhs_norm2)
        << ", cg_tol = " << tol << std::endl;
    throw ConvergenceError("Conjugate gradient has not converged." + err.str());
  }
  return x_k;
}

bool check_symmetry(const Matrix& eps, Real rel_tol) {
  Real a = 0, n = 0;
  for (Dim_t j = 0; j < eps.cols; ++j)
    for (Dim_t i = 0; i < eps.rows; ++i) {
      const Real d = eps(i, j) - eps(j, i);
      a += d * d;
      n += eps(i, j) * eps(i, j);
    }
  a = std::sqrt(a);
  n = std::sqrt(n);
  return rel_tol >= a / n || rel_tol >= n;
}

namespace {

Real norm(const DeviceVector& v) {
  double s = 0;
  check(mugpu_dot(v.data(), v.data(), v.size(), &s));
  return std::sqrt(s);
}

void check_load_steps(Cell& cell, const LoadSteps_t& steps) {
  const auto shape = cell.get_strain_shape();
  for (const auto& d : steps) {
    if (!(d.rows == shape[0] && d.cols == shape[1])) {
      std::stringstream err;
      err << "Load increments need to be given in " << shape[0] << "×" << shape[1]
          << " matrices, but I got a " << d.rows << "×" << d.cols << " matrix";
      throw SolverError(err.str());
    }
    if (cell.get_formulation() == Formulation::small_strain && !check_symmetry(d))
      throw SolverError("all Δε must be symmetric!");
  }
}

Matrix identity_or_zero(Cell& cell) {
  const Dim_t d = cell.get_material_dim();
  Matrix m{d, d, std::vector<Real>((size_t)(d * d), 0.0)};
  if (cell.get_formulation() == Formulation::finite_strain)
    for (Dim_t i = 0; i < d; ++i) m.data[i * (d + 1)] = 1.0;
  return m;
}

struct ConvergenceState {
  Real incr_norm, grad_norm, stress_norm, newton_tol, equil_tol;
  std::string message{"Has not converged"};
  bool has_converged{false};
  // the convergence_test lambda of src/solver/solvers.cc:152-163
  bool test() {
    const bool incr_test = incr_norm / grad_norm <= newton_tol;
    const bool stress_test = stress_norm < equil_tol;
    if (incr_test) message = "Residual  tolerance reached";
    else if (stress_test) message = "Reached stress divergence tolerance";
    has_converged = incr_test || stress_test;
    return has_converged;
  }
};

void print_header(const std::string& kind, Cell& cell, SolverBase& solver, Real newton_tol,
                  const LoadSteps_t& steps) {
  const bool small = cell.get_formulation() == Formulation::small_strain;
  std::cout << kind << "-" << solver.get_name() << " for " << (small ? "small" : "finite")
            << " strain with" << std::endl
            << "newton_tol = " << newton_tol << ", cg_tol = " << solver.get_tol()
            << " maxiter = " << solver.get_maxiter() << " and Δ" << (small ? "ε" : "F") << " ="
            << std::endl;
  for (size_t s = 0; s < steps.size(); ++s) {
    std::cout << "Step " << s + 1 << ":" << std::endl;
    for (Dim_t i = 0; i < steps[s].rows; ++i) {
      for (Dim_t j = 0; j < steps[s].cols; ++j) std::cout << (j ? " " : "") << steps[s](i, j);
      std::cout << std::endl;
    }
  }
}

// shared body of newton_cg (src/solver/solvers.cc:44-217) and de_geus (:220-408)
std::vector<OptimizeResult> drive(bool geus, Cell& cell, const LoadSteps_t& load_steps,
                                  SolverBase& solver, Real newton_tol, Real equil_tol,
                                  Dim_t verbose, Uint newton_maxiter) {
  const Uint newton_bound = newton_maxiter ? newton_maxiter : solver.get_maxiter();
  const int64_t ndof = cell.get_nb_dof();
  DeviceVector incrF(ndof), rhs(ndof), DeltaF;
  if (geus) DeltaF.resize(ndof);
  solver.initialise();
  const bool small = cell.get_formulation() == Formulation::small_strain;
  const std::string strain_symb = small ? "ε" : "F";
  size_t count_width = 0;
  if (verbose > 0) {
    print_header(geus ? "de Geus" : "Newton", cell, solver, newton_tol, load_steps);
    count_width = size_t(std::log10(solver.get_maxiter())) + 1;
  }
  check_load_steps(cell, load_steps);
  cell.initialise();
  cell.set_uniform_strain(identity_or_zero(cell));
  std::vector<OptimizeResult> ret;
  const Dim_t d = cell.get_material_dim();
  Matrix previous{d, d, std::vector<Real>((size_t)(d * d), 0.0)};
  for (const auto& macro : load_steps) {
    Matrix delta{d, d, std::vector<Real>((size_t)(d * d))};
    for (size_t k = 0; k < delta.data.size(); ++k) delta.data[k] = macro.data[k] - previous.data[k];
    if (!geus) cell.add_uniform_strain(delta);
    ConvergenceState cs{2 * newton_tol, 1.0, 2 * equil_tol, newton_tol, equil_tol};
    Uint newt_iter = 0;
    for (; newt_iter < newton_bound && !cs.has_converged; ++newt_iter) {
      cell.evaluate_stress_tangent();
      if (geus && newt_iter == 0) {
        // DeltaF = uniform macro increment; rhs = -A DeltaF (solvers.cc:353-363)
        {
          std::vector<Real> h((size_t)ndof);
          for (int64_t p = 0; p < cell.size(); ++p)
            for (Dim_t k = 0; k < d * d; ++k) h[(size_t)(p * d * d + k)] = delta.data[k];
          DeltaF.upload(h.data());
        }
        cell.evaluate_projected_directional_stiffness(DeltaF, rhs);
        check(mugpu_scale_copy(-1.0, rhs.data(), rhs.data(), ndof));
        cs.stress_norm = norm(rhs);
        if (cs.test()) break;
        DeviceVector& x = solver.solve(rhs);
        check(mugpu_scale_copy(1.0, x.data(), incrF.data(), ndof));
        check(mugpu_axpby(1.0, DeltaF.data(), 1.0, cell.strain_data(), ndof));
      } else {
        check(mugpu_scale_copy(-1.0, cell.stress_data(), rhs.data(), ndof));
        cell.apply_projection(rhs);
        cs.stress_norm = norm(rhs);
        if (cs.test()) break;
        DeviceVector& x = solver.solve(rhs);
        check(mugpu_scale_copy(1.0, x.data(), incrF.data(), ndof));
      }
      check(mugpu_axpby(1.0, incrF.data(), 1.0, cell.strain_data(), ndof));
      cs.incr_norm = norm(incrF);
      double g2 = 0;
      check(mugpu_dot(cell.strain_data(), cell.strain_data(), ndof, &g2));
      cs.grad_norm = std::sqrt(g2);
      if (verbose > 0) {
        std::cout << "at Newton step " << std::setw(count_width) << newt_iter << ", |δ"
                  << strain_symb << "|/|Δ" << strain_symb << "| = " << std::setw(17)
                  << cs.incr_norm / cs.grad_norm << ", tol = " << newton_tol << std::endl;
      }
      cs.test();
    }
    previous = macro;
    OptimizeResult r;
    r.grad.capture(cell.strain_data(), ndof);
    r.stress.capture(cell.stress_data(), ndof);
    r.success = cs.test();
    r.status = Int(r.success);
    r.message = cs.message;
    r.nb_it = newt_iter;
    r.nb_fev = solver.get_counter();
    ret.push_back(std::move(r));
    cell.save_history_variables();
  }
  return ret;
}

}  // namespace

std::vector<OptimizeResult> newton_cg(Cell& cell, const LoadSteps_t& load_steps,
                                      SolverBase& solver, Real newton_tol, Real equil_tol,
                                      Dim_t verbose, Uint newton_maxiter) {
  return drive(false, cell, load_steps, solver, newton_tol, equil_tol, verbose, newton_maxiter);
}

std::vector<OptimizeResult> de_geus(Cell& cell, const LoadSteps_t& load_steps,
                                    SolverBase& solver, Real newton_tol, Real equil_tol,
                                    Dim_t verbose, Uint newton_maxiter) {
  return drive(true, cell, load_steps, solver, newton_tol, equil_tol, verbose, newton_maxiter);
}

}  // namespace muSpectre
