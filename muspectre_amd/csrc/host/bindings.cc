// pybind11 module _muSpectre with the reference's Python surface for the
// Newton-CG path (language_bindings/python/bind_py_{module,common,cell,
// material,solvers,fftengine,projections}.cc), backed by the HIP host layer.
//
// Array conventions are the reference's: fields are (nb_components x
// nb_pixels) arrays in column-major (Fortran) order, i.e. AoS memory
// (bind_py_cell.cc:133-140); complex spectra are (nb_components x
// nb_fourier_points).
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstring>
#include <sstream>

#include "muspectre_host.hh"

namespace py = pybind11;
using namespace pybind11::literals;
using namespace muSpectre;

namespace {

using farr = py::array_t<double, py::array::f_style | py::array::forcecast>;
using carr = py::array_t<std::complex<double>, py::array::f_style | py::array::forcecast>;
using rarr = py::array_t<double, py::array::c_style | py::array::forcecast>;

std::string shape_error(int64_t rows, int64_t cols, const py::array& a) {
  std::stringstream err;
  err << "need array of shape (" << rows << ", " << cols << ") but got (";
  for (py::ssize_t i = 0; i < a.ndim(); ++i) err << (i ? ", " : "") << a.shape(i);
  err << ").";
  return err.str();
}

std::vector<Real> field_from(const farr& a, int64_t rows, int64_t cols) {
  if (a.ndim() != 2 || a.shape(0) != rows || a.shape(1) != cols)
    throw std::runtime_error(shape_error(rows, cols, a));
  return std::vector<Real>(a.data(), a.data() + rows * cols);
}

py::array field_to(const std::vector<Real>& v, int64_t rows, int64_t cols) {
  py::array_t<double, py::array::f_style> out({(py::ssize_t)rows, (py::ssize_t)cols});
  std::copy(v.begin(), v.end(), out.mutable_data());
  return std::move(out);
}

// a (dim x dim) numpy array -> column-major Matrix (Eigen::MatrixXd semantics)
Matrix matrix_from(const rarr& a) {
  if (a.ndim() != 2) throw SolverError("load increments must be 2D arrays");
  Matrix m{(Dim_t)a.shape(0), (Dim_t)a.shape(1), {}};
  m.data.resize((size_t)(m.rows * m.cols));
  for (Dim_t i = 0; i < m.rows; ++i)
    for (Dim_t j = 0; j < m.cols; ++j) m.data[i + m.rows * j] = a.at(i, j);
  return m;
}

LoadSteps_t steps_from(const py::sequence& seq) {
  LoadSteps_t steps;
  for (auto item : seq) steps.push_back(matrix_from(item.cast<rarr>()));
  return steps;
}

Precision precision_from(const py::object& p) {
  if (p.is_none()) return Precision::fp64;
  if (py::isinstance<py::str>(p)) {
    const std::string s = p.cast<std::string>();
    if (s == "fp64" || s == "float64" || s == "double") return Precision::fp64;
    if (s == "fp32" || s == "float32" || s == "mixed") return Precision::fp32;
    throw std::runtime_error("unknown precision '" + s + "'");
  }
  return p.cast<Precision>();
}

template <Dim_t Dim>
int64_t pixel_index(Cell& cell, const Ccoord_t<Dim>& pix);

// pixels of a material of no cell are numbered in the order they come
template <Dim_t Dim>
int64_t material_pixel(MaterialBase& mat, const Ccoord_t<Dim>& pix) {
  return mat.has_cell() ? pixel_index<Dim>(mat.get_cell(), pix) : (int64_t)mat.size();
}

template <Dim_t Dim>
int64_t pixel_index(Cell& cell, const Ccoord_t<Dim>& pix) {
  const auto& res = cell.get_resolutions();
  int64_t i = 0;
  for (int d = 0; d < Dim; ++d) {
    if (pix[d] < 0 || pix[d] >= res[d]) throw std::runtime_error("pixel outside the cell");
    i = i * res[d] + pix[d];
  }
  return i;
}

// ---- per-family material classes (types for the python names) ----------------
template <class Solver>
void add_eigen_solver(py::module& m, const char* name) {
  py::class_<Solver, SolverEigenBase>(m, name)
      .def(py::init<Cell&, Real, Uint, bool>(), "cell"_a, "tol"_a, "maxiter"_a, "verbose"_a = false,
           py::keep_alive<1, 2>())
      .def("name", &Solver::get_name);
}

// MaterialBase<dim, dim> (the python Material_2d / Material_3d,
// bind_py_material.cc:158-180), base of the per-family classes
template <Dim_t Dim>
struct MaterialDim : MaterialBase {
  using MaterialBase::MaterialBase;
};

template <Dim_t Dim, int Fam>
struct MaterialT : MaterialDim<Dim> {
  MaterialT(Cell& c, const std::string& n, Real E, Real nu, Real tau, Real H, int fam = Fam)
      : MaterialDim<Dim>(c, n, fam, E, nu, tau, H) {}
  MaterialT(Cell& c, const std::string& n, const Matrix& C_voigt)
      : MaterialDim<Dim>(c, n, C_voigt) {}
  explicit MaterialT(const std::string& n) : MaterialDim<Dim>(n, Fam, Dim) {}
};

// ---- a material's internal field collection (bind_py_field_collection.cc) ----
// LocalFieldCollection of a material: names, typed access; every field of
// this build's materials holds Real values, so the Int, Uint and Complex
// getters raise as the reference's get_typed_field does for a field of
// another type (field_collection_base.hh:282-293)
// (one C++ type per dimension: the 2D and 3D python classes are distinct)
template <Dim_t Dim>
struct FieldCollectionView {
  MaterialBase* mat;
};
template <Dim_t Dim>
struct FieldView {
  MaterialBase* mat;
  std::string name;
};
template <Dim_t Dim>
struct StateFieldView {
  MaterialBase* mat;
  std::string prefix;
};

template <Dim_t Dim>
py::array field_array(const FieldView<Dim>& f) {
  Dim_t nb = 0;
  auto v = f.mat->field_values(f.name, nb);
  return field_to(v, nb, (int64_t)f.mat->size());
}

[[noreturn]] void wrong_type(MaterialBase& m, const std::string& name, const char* type) {
  (void)m.field_values(name, *std::make_unique<Dim_t>());  // raises for unknown names
  throw std::runtime_error("Field '" + name + "' is of type d, but should be of type " + type);
}

template <Dim_t Dim>
void bind_field_collection(py::module& m) {
  const std::string d = std::to_string(Dim);
  py::class_<FieldView<Dim>>(m, ("LocalFieldReal_" + d).c_str())
      .def_property_readonly("name", [](const FieldView<Dim>& f) { return f.name; }, "field name")
      .def_property_readonly(
          "collection",
          py::cpp_function([](const FieldView<Dim>& f) { return FieldCollectionView<Dim>{f.mat}; },
                           py::keep_alive<0, 1>()),
          "Collection containing this field")
      .def_property_readonly("nb_components",
                             [](const FieldView<Dim>& f) {
                               Dim_t nb = 0;
                               f.mat->field_values(f.name, nb);
                               return nb;
                             },
                             "number of scalars stored per pixel in this field")
      .def_property_readonly("stored_type", [](const FieldView<Dim>&) { return std::string("d"); },
                             "fundamental type of scalars stored in this field")
      .def_property_readonly("size", [](const FieldView<Dim>& f) { return f.mat->size(); },
                             "number of pixels in this field")
      .def("set_zero",
           [](FieldView<Dim>& f) {
             Dim_t nb = 0;
             auto v = f.mat->field_values(f.name, nb);
             f.mat->set_field_values(f.name, std::vector<Real>(v.size(), 0.0));
           },
           "Set all components in the field to zero")
      .def_property(
          "array", [](const FieldView<Dim>& f) { return field_array(f); },
          [](FieldView<Dim>& f, farr a) {
            Dim_t nb = 0;
            f.mat->field_values(f.name, nb);
            f.mat->set_field_values(f.name, field_from(a, nb, (int64_t)f.mat->size()));
          },
          "array of stored data")
      .def_property(
          "vector",
          [](const FieldView<Dim>& f) {
            Dim_t nb = 0;
            auto v = f.mat->field_values(f.name, nb);
            return field_to(v, (int64_t)v.size(), 1);
          },
          [](FieldView<Dim>& f, farr a) {
            Dim_t nb = 0;
            auto v = f.mat->field_values(f.name, nb);
            f.mat->set_field_values(f.name, field_from(a, (int64_t)v.size(), 1));
          },
          "flattened array of stored data");
  py::class_<StateFieldView<Dim>>(m, ("LocalStateFieldReal_" + d).c_str())
      .def_property_readonly("prefix", [](const StateFieldView<Dim>& f) { return f.prefix; },
                             "state field prefix")
      .def_property_readonly(
          "collection",
          py::cpp_function([](const StateFieldView<Dim>& f) { return FieldCollectionView<Dim>{f.mat}; },
                           py::keep_alive<0, 1>()),
          "Collection containing this field")
      .def_property_readonly("nb_memory", [](const StateFieldView<Dim>&) { return 1; },
                             "number of old states stored")
      .def_property_readonly("stored_type", [](const StateFieldView<Dim>&) { return std::string("d"); },
                             "fundamental type of scalars stored in this field")
      // statefield.hh:245-269: current is sub-field indices[0], old indices[1]
      .def("get_current_field",
           [](const StateFieldView<Dim>& f) {
             const int cur = f.mat->current_sub_field(f.prefix);
             return FieldView<Dim>{f.mat, f.prefix + ", sub_field index " + std::to_string(cur)};
           },
           "returns the current field value", py::keep_alive<0, 1>())
      .def("get_old_field",
           [](const StateFieldView<Dim>& f, size_t nb_steps_ago) {
             if (nb_steps_ago != 1) throw py::index_error("array::at");
             const int cur = f.mat->current_sub_field(f.prefix);
             return FieldView<Dim>{f.mat, f.prefix + ", sub_field index " + std::to_string(1 - cur)};
           },
           "nb_steps_ago"_a = 1, "returns the value this field held 'nb_steps_ago' steps ago",
           py::keep_alive<0, 1>());
  auto get_field = [](FieldCollectionView<Dim>& c, const std::string& n) {
    Dim_t nb = 0;
    c.mat->field_values(n, nb);  // raises for a name the collection does not hold
    return FieldView<Dim>{c.mat, n};
  };
  auto get_state = [](FieldCollectionView<Dim>& c, const std::string& n) {
    const auto names = c.mat->statefield_names();
    if (std::find(names.begin(), names.end(), n) == names.end())
      throw std::runtime_error("The state field named '" + n + "' does not exist in the material '" +
                               c.mat->get_name() + "'");
    return StateFieldView<Dim>{c.mat, n};
  };
  py::class_<FieldCollectionView<Dim>>(m, ("_LocalFieldCollection_" + d + "d").c_str())
      .def("get_real_field", get_field, "unique_name"_a, py::keep_alive<0, 1>())
      .def("get_int_field",
           [](FieldCollectionView<Dim>& c, const std::string& n) -> FieldView<Dim> { wrong_type(*c.mat, n, "i"); },
           "unique_name"_a)
      .def("get_uint_field",
           [](FieldCollectionView<Dim>& c, const std::string& n) -> FieldView<Dim> { wrong_type(*c.mat, n, "j"); },
           "unique_name"_a)
      .def("get_complex_field",
           [](FieldCollectionView<Dim>& c, const std::string& n) -> FieldView<Dim> {
             wrong_type(*c.mat, n, "St7complexIdE");
           },
           "unique_name"_a)
      .def("get_real_statefield", get_state, "unique_name"_a, py::keep_alive<0, 1>())
      .def("get_int_statefield",
           [](FieldCollectionView<Dim>& c, const std::string& n) -> StateFieldView<Dim> {
             throw std::runtime_error("State field '" + n + "' does not hold values of type i");
           },
           "unique_name"_a)
      .def("get_uint_statefield",
           [](FieldCollectionView<Dim>& c, const std::string& n) -> StateFieldView<Dim> {
             throw std::runtime_error("State field '" + n + "' does not hold values of type j");
           },
           "unique_name"_a)
      .def("get_complex_statefield",
           [](FieldCollectionView<Dim>& c, const std::string& n) -> StateFieldView<Dim> {
             throw std::runtime_error("State field '" + n +
                                      "' does not hold values of type St7complexIdE");
           },
           "unique_name"_a)
      .def_property_readonly("field_names",
                             [](FieldCollectionView<Dim>& c) { return c.mat->field_names(); },
                             "returns the names of all fields in this collection")
      .def_property_readonly("statefield_names",
                             [](FieldCollectionView<Dim>& c) { return c.mat->statefield_names(); },
                             "returns the names of all state fields in this collection");
  py::class_<MaterialDim<Dim>, MaterialBase>(m, ("Material_" + d + "d").c_str())
      .def_property_readonly(
          "collection",
          py::cpp_function([](MaterialDim<Dim>& mat) { return FieldCollectionView<Dim>{&mat}; },
                           py::keep_alive<0, 1>()),
          "returns the field collection containing internal fields of this material");
}

template <Dim_t Dim, int Fam>
MaterialT<Dim, Fam>& attach(CellBase<Dim, Dim>& sys, MaterialT<Dim, Fam>* m) {
  return static_cast<MaterialT<Dim, Fam>&>(sys.add_material(std::unique_ptr<MaterialBase>(m)));
}

template <Dim_t Dim, int Fam>
void bind_material(py::module& m, const std::string& stem) {
  using Mat = MaterialT<Dim, Fam>;
  using Sys = CellBase<Dim, Dim>;
  const std::string name = stem + "_" + std::to_string(Dim) + "d";
  py::class_<Mat, MaterialDim<Dim>> cls(m, name.c_str());
  if (Fam == MUGPU_LE3 || Fam == MUGPU_LE4) {
    // bind_py_material.cc:115,141: a material of no cell
    cls.def(py::init<const std::string&>(), "name"_a);
  }
  if (Fam == MUGPU_LE1 || Fam == MUGPU_LE2) {
    // bind_py_material.cc:66-71
    cls.def_static(
        "make",
        [](Sys& sys, std::string n, Real E, Real nu) -> Mat& {
          return attach<Dim, Fam>(sys, new Mat(sys, n, E, nu, 0, 0));
        },
        "cell"_a, "name"_a, "Young"_a, "Poisson"_a, py::return_value_policy::reference,
        py::keep_alive<1, 0>());
  } else if (Fam == MUGPU_LEG) {
    // MaterialLinearElasticGeneric(name, C_voigt)
    // (material_linear_elastic_generic.cc:40-65; not bound in the reference)
    cls.def_static(
        "make",
        [](Sys& sys, std::string n, rarr Cv) -> Mat& {
          if (Cv.ndim() != 2) throw std::runtime_error("C_voigt must be a matrix");
          Matrix C;
          C.rows = (Dim_t)Cv.shape(0);
          C.cols = (Dim_t)Cv.shape(1);
          C.data.resize((size_t)C.rows * C.cols);
          for (Dim_t j = 0; j < C.cols; ++j)
            for (Dim_t i = 0; i < C.rows; ++i) C.data[i + C.rows * j] = Cv.at(i, j);
          return attach<Dim, Fam>(sys, new Mat(sys, n, C));
        },
        "cell"_a, "name"_a, "C_voigt"_a, py::return_value_policy::reference,
        py::keep_alive<1, 0>());
  } else if (Fam == MUGPU_HEP1) {
    // material_hyper_elasto_plastic1.cc:40-65 (not bound in the reference);
    // consistent_tangent=True keeps the stress update and replaces the
    // coded tangent by dP/dF
    cls.def_static(
        "make",
        [](Sys& sys, std::string n, Real E, Real nu, Real tau, Real H, bool ct) -> Mat& {
          return attach<Dim, Fam>(sys, new Mat(sys, n, E, nu, tau, H, ct ? MUGPU_HEP1_CT : MUGPU_HEP1));
        },
        "cell"_a, "name"_a, "Young"_a, "Poisson"_a, "tau_y0"_a, "H"_a,
        "consistent_tangent"_a = false, py::return_value_policy::reference,
        py::keep_alive<1, 0>());
  } else {
    cls.def_static(
        "make",
        [](Sys& sys, std::string n) -> Mat& { return attach<Dim, Fam>(sys, new Mat(sys, n, 0, 0, 0, 0)); },
        "cell"_a, "name"_a, py::return_value_policy::reference, py::keep_alive<1, 0>());
  }
  if (Fam == MUGPU_LE2) {
    // bind_py_material.cc:95-100: eigenstrain as a dim x dim matrix
    cls.def(
        "add_pixel",
        [](Mat& mat, Ccoord_t<Dim> pix, rarr eig) {
          if (eig.ndim() != 2 || eig.shape(0) != Dim || eig.shape(1) != Dim)
            throw std::runtime_error("eigenstrain must be a dim x dim array");
          std::vector<Real> p(Dim * Dim);
          for (int i = 0; i < Dim; ++i)
            for (int j = 0; j < Dim; ++j) p[i + Dim * j] = eig.at(i, j);
          mat.add_pixel_index_with(material_pixel<Dim>(mat, pix), p);
        },
        "pixel"_a, "eigenstrain"_a);
  } else if (Fam == MUGPU_LE3 || Fam == MUGPU_LE4) {
    cls.def(
        "add_pixel",
        [](Mat& mat, Ccoord_t<Dim> pix, Real E, Real nu) {
          mat.add_pixel_index_with(material_pixel<Dim>(mat, pix), {E, nu});
        },
        "pixel"_a, "Young"_a, "Poisson"_a);
  } else {
    cls.def(
        "add_pixel",
        [](Mat& mat, Ccoord_t<Dim> pix) {
          mat.add_pixel_index(material_pixel<Dim>(mat, pix));
        },
        "pixel"_a);
  }
  // vectorised pixel assignment by linear index (no reference counterpart;
  // adding millions of pixels one call at a time dominates set-up time)
  cls.def(
      "add_pixels_by_index",
      [](Mat& mat, py::array_t<int64_t, py::array::c_style | py::array::forcecast> idx,
         py::object params) {
        const int64_t n = idx.size();
        const int64_t* p = idx.data();
        if (params.is_none()) {
          for (int64_t i = 0; i < n; ++i) mat.add_pixel_index(p[i]);
          return;
        }
        rarr pa = params.cast<rarr>();
        const int64_t per = n ? pa.size() / n : 0;
        if (per * n != pa.size()) throw std::runtime_error("params must have one row per pixel");
        const double* q = pa.data();
        for (int64_t i = 0; i < n; ++i)
          mat.add_pixel_index_with(p[i], std::vector<Real>(q + i * per, q + (i + 1) * per));
      },
      "indices"_a, "params"_a = py::none());
  cls.def("size", &Mat::size);
}

template <Dim_t Dim>
struct PixelIterator {
  std::vector<Dim_t> res;
  int64_t i{0}, n{0};
};

template <Dim_t Dim>
void bind_cell(py::module& m) {
  using Sys = CellBase<Dim, Dim>;
  const std::string name = "CellBase" + std::to_string(Dim) + "d";
  const std::string itname = "PixelIterator" + std::to_string(Dim) + "d";
  py::class_<PixelIterator<Dim>>(m, itname.c_str())
      .def("__iter__", [](PixelIterator<Dim>& it) -> PixelIterator<Dim>& { return it; })
      .def("__next__", [](PixelIterator<Dim>& it) {
        if (it.i >= it.n) throw py::stop_iteration();
        Ccoord_t<Dim> c;
        int64_t r = it.i++;
        for (int d = Dim - 1; d >= 0; --d) {
          c[d] = (Dim_t)(r % it.res[d]);
          r /= it.res[d];
        }
        return c;
      });
  const int nc = Dim * Dim;
  py::class_<Sys, Cell>(m, name.c_str())
      .def("__len__", [](Sys& s) { return s.size(); })
      .def("__iter__", [](Sys& s) { return PixelIterator<Dim>{s.get_resolutions(), 0, s.size()}; })
      .def("initialise", [](Sys& s, FFT_PlanFlags f) { s.initialise(f); }, "flags"_a = FFT_PlanFlags::estimate)
      .def(
          "directional_stiffness",
          [nc](Sys& s, farr v) {
            auto h = field_from(v, nc, s.size());
            if (!s.is_initialised()) s.initialise();
            DeviceVector in(s.get_nb_dof()), out(s.get_nb_dof());
            in.upload(h.data());
            s.evaluate_projected_directional_stiffness(in, out);
            return field_to(out.to_host(), nc, s.size());
          },
          "δF"_a)
      .def(
          "project",
          [nc](Sys& s, farr v) {
            auto h = field_from(v, nc, s.size());
            if (!s.is_initialised()) s.initialise();
            DeviceVector buf(s.get_nb_dof());
            buf.upload(h.data());
            s.apply_projection(buf);
            return field_to(buf.to_host(), nc, s.size());
          },
          "field"_a)
      .def("get_strain", [nc](Sys& s) { return field_to(s.get_strain(), nc, s.size()); })
      .def("get_stress", [nc](Sys& s) { return field_to(s.get_stress(), nc, s.size()); })
      .def("get_tangent", [nc](Sys& s) { return field_to(s.get_tangent(), nc * nc, s.size()); })
      .def_property_readonly("size", [](Sys& s) { return s.size(); })
      .def(
          "evaluate_stress_tangent",
          [nc](Sys& s, farr v) {
            auto h = field_from(v, nc, s.size());
            s.initialise();
            s.set_strain(h);
            s.evaluate_stress_tangent();
          },
          "strain"_a)
      .def("save_history_variables", &Sys::save_history_variables)
      .def("get_projection",
           [](Sys& s) { return CellProjection<Dim>(s); },
           "the projection of the cell (CellBase::get_projection)", py::keep_alive<0, 1>())
      .def("get_subdomain_resolutions", &Sys::get_subdomain_resolutions)
      .def("get_subdomain_locations", &Sys::get_subdomain_locations)
      .def("get_domain_resolutions", &Sys::get_domain_resolutions)
      // bind_py_cell.cc:202 binds get_domain_lengths to the resolutions
      .def("get_domain_lengths", &Sys::get_domain_resolutions)
      .def("get_precision", &Sys::get_precision);
}

template <Dim_t Dim>
void bind_cell_factory(py::module& m) {
  using Ccoord = Ccoord_t<Dim>;
  using Rcoord = Rcoord_t<Dim>;
  Rcoord ones;
  ones.fill(1.0);
  // bind_py_cell.cc:86-98, with two keyword-only extensions: the operator
  // precision and the even-grid Nyquist policy
  m.def(
      "CellFactory",
      [](Ccoord res, Rcoord lens, Formulation form, py::object precision, bool allow_even) {
        return make_cell<Dim, Dim>(res, lens, form, precision_from(precision), allow_even);
      },
      "resolutions"_a, "lengths"_a = ones, "formulation"_a = Formulation::finite_strain,
      py::kw_only(), "precision"_a = py::none(), "allow_even"_a = false);
}

template <Dim_t Dim>
void bind_engine(py::module& m, const std::string& name) {
  using Eng = HIPEngine<Dim>;
  py::class_<Eng>(m, name.c_str())
      .def(py::init<Ccoord_t<Dim>, Dim_t>(), "resolutions"_a, "nb_components"_a)
      .def(
          "fft",
          [](Eng& eng, farr v) {
            const int64_t nc = eng.get_nb_components();
            auto h = field_from(v, nc, eng.size());
            DeviceVector in(nc * eng.size());
            in.upload(h.data());
            auto& w = eng.fft(in);
            auto spec = w.to_host();
            py::array_t<std::complex<double>, py::array::f_style> out(
                {(py::ssize_t)nc, (py::ssize_t)eng.fourier_size()});
            std::memcpy(out.mutable_data(), spec.data(), sizeof(double) * spec.size());
            return out;
          },
          "array"_a)
      .def(
          "ifft",
          [](Eng& eng, carr v) {
            const int64_t nc = eng.get_nb_components();
            if (v.ndim() != 2 || v.shape(0) != nc || v.shape(1) != eng.fourier_size())
              throw std::runtime_error(shape_error(nc, eng.fourier_size(), v));
            if (!eng.is_initialised()) throw std::runtime_error("ifft plan not initialised");
            eng.get_work_space().upload(reinterpret_cast<const double*>(v.data()));
            DeviceVector out(nc * eng.size());
            eng.ifft(out);
            return field_to(out.to_host(), nc, eng.size());
          },
          "array"_a)
      .def("initialise", &Eng::initialise, "flags"_a = FFT_PlanFlags::estimate)
      .def("normalisation", &Eng::normalisation)
      .def("get_subdomain_resolutions", &Eng::get_subdomain_resolutions)
      .def("get_subdomain_locations", &Eng::get_subdomain_locations)
      .def("get_fourier_resolutions", &Eng::get_fourier_resolutions)
      .def("get_fourier_locations", &Eng::get_fourier_locations)
      .def("get_domain_resolutions", &Eng::get_domain_resolutions);
}

template <class Proj, Dim_t Dim>
void bind_projection_methods(py::class_<Proj>& cls) {
  const int nc = Dim * Dim;
  cls.def("initialise", &Proj::initialise, "flags"_a = FFT_PlanFlags::estimate,
           "initialises the fft engine (plan the transform)")
      .def("apply_projection",
           [nc](Proj& p, farr v) {
             auto h = field_from(v, nc, p.size());
             DeviceVector buf(nc * p.size());
             buf.upload(h.data());
             p.apply_projection(buf);
             return field_to(buf.to_host(), nc, p.size());
           })
      .def("get_operator",
           [](Proj& p) {
             auto op = p.get_operator();
             const int64_t k = p.fourier_size();
             return field_to(op, (int64_t)op.size() / k, k);
           })
      .def("get_formulation", &Proj::get_formulation,
           "return a Formulation enum indicating whether the projection is small or finite strain")
      .def("get_subdomain_resolutions", &Proj::get_subdomain_resolutions)
      .def("get_subdomain_locations", &Proj::get_subdomain_locations)
      .def("get_domain_resolutions", &Proj::get_domain_resolutions)
      // bind_py_projections.cc:181 binds get_domain_lengths to the resolutions
      .def("get_domain_lengths", &Proj::get_domain_resolutions);
}

template <class Proj, Dim_t Dim>
void bind_projection(py::module& m, const std::string& stem) {
  const std::string name = stem + "_" + std::to_string(Dim) + "d";
  py::class_<Proj> cls(m, name.c_str());
  cls.def(py::init([](Ccoord_t<Dim> res, Rcoord_t<Dim> lengths, const std::string& fft) {
            if (fft != "fftw" && fft != "hip")
              throw std::runtime_error("Unknown FFT engine '" + fft + "' specified.");
            std::unique_ptr<HIPEngine<Dim>> eng(new HIPEngine<Dim>(res, Dim * Dim));
            return std::unique_ptr<Proj>(new Proj(std::move(eng), lengths));
          }),
          "resolutions"_a, "lengths"_a, "fft"_a = "fftw");
  bind_projection_methods<Proj, Dim>(cls);
}

template <Dim_t Dim>
void bind_common_helpers(py::module& m) {
  using Ccoord = Ccoord_t<Dim>;
  m.def(
      "get_hermitian_sizes",
      [](Ccoord s) {
        s[Dim - 1] = s[Dim - 1] / 2 + 1;
        return s;
      },
      "full_sizes"_a, "return the hermitian sizes corresponding to the true sizes");
  m.def(
      "get_domain_ccoord",
      [](Ccoord res, Dim_t index) {
        Ccoord c;
        for (int d = Dim - 1; d >= 0; --d) {
          c[d] = index % res[d];
          index /= res[d];
        }
        return c;
      },
      "resolutions"_a, "i"_a);
  m.def(
      "get_domain_index",
      [](Ccoord sizes, Ccoord c) {
        Dim_t i = 0;
        for (int d = 0; d < Dim; ++d) i = i * sizes[d] + c[d];
        return i;
      },
      "sizes"_a, "ccoord"_a);
  m.def(("get_" + std::to_string(Dim) + "d_cube").c_str(), [](Dim_t s) {
    Ccoord c;
    c.fill(s);
    return c;
  }, "size"_a);
  m.def(("get_" + std::to_string(Dim) + "d_cube").c_str(), [](Real s) {
    Rcoord_t<Dim> c;
    c.fill(s);
    return c;
  }, "size"_a);
}

py::object solve_driver(bool geus, Cell& cell, py::object steps_obj, SolverBase& solver,
                        Real newton_tol, Real equil_tol, Dim_t verbose, Uint newton_maxiter) {
  const bool is_list = py::isinstance<py::list>(steps_obj) || py::isinstance<py::tuple>(steps_obj);
  LoadSteps_t steps = is_list ? steps_from(steps_obj.cast<py::sequence>())
                              : LoadSteps_t{matrix_from(steps_obj.cast<rarr>())};
  std::vector<OptimizeResult> res;
  {
    res = geus ? de_geus(cell, steps, solver, newton_tol, equil_tol, verbose, newton_maxiter)
               : newton_cg(cell, steps, solver, newton_tol, equil_tol, verbose, newton_maxiter);
  }
  if (!is_list) return py::cast(std::move(res.front()));
  py::list out;
  for (auto& r : res) out.append(py::cast(std::move(r)));
  return std::move(out);
}

}  // namespace

PYBIND11_MODULE(_muSpectre, mod) {
  mod.doc() = "Python bindings to the µSpectre Newton-CG path on MI355X (HIP)";

  // ---- common (bind_py_common.cc) ----
  py::enum_<Formulation>(mod, "Formulation")
      .value("finite_strain", Formulation::finite_strain)
      .value("small_strain", Formulation::small_strain);
  py::enum_<StressMeasure>(mod, "StressMeasure")
      .value("Cauchy", StressMeasure::Cauchy)
      .value("PK1", StressMeasure::PK1)
      .value("PK2", StressMeasure::PK2)
      .value("Kirchhoff", StressMeasure::Kirchhoff)
      .value("Biot", StressMeasure::Biot)
      .value("Mandel", StressMeasure::Mandel)
      .value("no_stress_", StressMeasure::no_stress_);
  py::enum_<StrainMeasure>(mod, "StrainMeasure")
      .value("Gradient", StrainMeasure::Gradient)
      .value("Infinitesimal", StrainMeasure::Infinitesimal)
      .value("GreenLagrange", StrainMeasure::GreenLagrange)
      .value("Biot", StrainMeasure::Biot)
      .value("Log", StrainMeasure::Log)
      .value("Almansi", StrainMeasure::Almansi)
      .value("RCauchyGreen", StrainMeasure::RCauchyGreen)
      .value("LCauchyGreen", StrainMeasure::LCauchyGreen)
      .value("no_strain_", StrainMeasure::no_strain_);
  py::enum_<FFT_PlanFlags>(mod, "FFT_PlanFlags")
      .value("estimate", FFT_PlanFlags::estimate)
      .value("measure", FFT_PlanFlags::measure)
      .value("patient", FFT_PlanFlags::patient);
  py::enum_<Precision>(mod, "Precision").value("fp64", Precision::fp64).value("fp32", Precision::fp32);
  // the notice executables print at start-up (common.hh:130-134, bound by
  // bind_py_common.cc:158 with these argument names)
  mod.def(
      "banner",
      [](const std::string& name, Uint year, const std::string& holder) {
        py::print("");
        py::print("µSpectre " + name);
        py::print("Copyright © " + std::to_string(year) + "  " + holder);
        py::print("Free software without any warranty; see the licence file for the terms "
                  "of use and redistribution.");
        py::print("");
      },
      "name"_a, "year"_a, "copyright_holder"_a);
  bind_common_helpers<2>(mod);
  bind_common_helpers<3>(mod);

  py::register_exception<SolverError>(mod, "SolverError", PyExc_RuntimeError);
  py::register_exception<ConvergenceError>(mod, "ConvergenceError", PyExc_RuntimeError);
  py::register_exception<ProjectionError>(mod, "ProjectionError", PyExc_RuntimeError);
  py::register_exception<DeviceError>(mod, "DeviceError", PyExc_RuntimeError);

  // ---- cell (bind_py_cell.cc) ----
  bind_cell_factory<2>(mod);
  bind_cell_factory<3>(mod);
  py::class_<Cell>(mod, "Cell").def(
      "get_globalised_internal_real_array",
      [](Cell& c, const std::string& name) {
        Dim_t nb = 0;
        auto v = c.get_globalised_internal_real_array(name, nb);
        return field_to(v, nb, c.size());
      },
      "unique_name"_a);
  bind_cell<2>(mod);
  bind_cell<3>(mod);

  // ---- materials (bind_py_material.cc) ----
  auto material = mod.def_submodule("material");
  material.doc() = "bindings for constitutive laws";
  py::class_<MaterialBase>(material, "MaterialBase")
      .def("get_name", &MaterialBase::get_name)
      .def("size", &MaterialBase::size);
  bind_field_collection<2>(material);
  bind_field_collection<3>(material);
  bind_material<2, MUGPU_LE1>(material, "MaterialLinearElastic1");
  bind_material<3, MUGPU_LE1>(material, "MaterialLinearElastic1");
  bind_material<2, MUGPU_LE2>(material, "MaterialLinearElastic2");
  bind_material<3, MUGPU_LE2>(material, "MaterialLinearElastic2");
  bind_material<2, MUGPU_LE3>(material, "MaterialLinearElastic3");
  bind_material<3, MUGPU_LE3>(material, "MaterialLinearElastic3");
  bind_material<2, MUGPU_LE4>(material, "MaterialLinearElastic4");
  bind_material<3, MUGPU_LE4>(material, "MaterialLinearElastic4");
  bind_material<2, MUGPU_LEG>(material, "MaterialLinearElasticGeneric");
  bind_material<3, MUGPU_LEG>(material, "MaterialLinearElasticGeneric");
  bind_material<2, MUGPU_HEP1>(material, "MaterialHyperElastoPlastic1");
  bind_material<3, MUGPU_HEP1>(material, "MaterialHyperElastoPlastic1");

  // ---- solvers (bind_py_solvers.cc) ----
  py::class_<OptimizeResult>(mod, "OptimizeResult")
      .def(py::init<>())
      .def_property("grad",
                    [](const OptimizeResult& r) { return field_to(r.grad.host(), r.grad.size(), 1); },
                    [](OptimizeResult& r, farr a) { r.grad.assign(a.data(), (int64_t)a.size()); })
      .def_property("stress",
                    [](const OptimizeResult& r) { return field_to(r.stress.host(), r.stress.size(), 1); },
                    [](OptimizeResult& r, farr a) { r.stress.assign(a.data(), (int64_t)a.size()); })
      .def_readwrite("success", &OptimizeResult::success)
      .def_readwrite("status", &OptimizeResult::status)
      .def_readwrite("message", &OptimizeResult::message)
      .def_readwrite("nb_it", &OptimizeResult::nb_it)
      .def_readwrite("nb_fev", &OptimizeResult::nb_fev);
  auto solvers = mod.def_submodule("solvers");
  solvers.doc() = "bindings for solvers";
  py::class_<SolverBase>(solvers, "SolverBase")
      .def("get_counter", &SolverBase::get_counter)
      .def("reset_counter", &SolverBase::reset_counter)
      .def("has_converged", &SolverBase::has_converged)
      .def("get_maxiter", &SolverBase::get_maxiter)
      .def("get_tol", &SolverBase::get_tol);
  py::class_<SolverCG, SolverBase>(solvers, "SolverCG")
      .def(py::init<Cell&, Real, Uint, bool>(), "cell"_a, "tol"_a, "maxiter"_a, "verbose"_a = false,
           py::keep_alive<1, 2>())
      .def("name", &SolverCG::get_name)
      .def("set_inner_tol", &SolverCG::set_inner_tol, "tol"_a)
      .def("get_refinements", &SolverCG::get_refinements);
  // Eigen's iterative solvers (bind_py_solvers.cc:52-72)
  py::class_<SolverEigenBase, SolverBase>(solvers, "SolverEigenBase");
  add_eigen_solver<SolverCGEigen>(solvers, "SolverCGEigen");
  add_eigen_solver<SolverGMRESEigen>(solvers, "SolverGMRESEigen");
  add_eigen_solver<SolverBiCGSTABEigen>(solvers, "SolverBiCGSTABEigen");
  add_eigen_solver<SolverDGMRESEigen>(solvers, "SolverDGMRESEigen");
  add_eigen_solver<SolverMINRESEigen>(solvers, "SolverMINRESEigen");
  // single-step overload first: kwarg equil_tol (bind_py_solvers.cc:82-94),
  // list overload: kwarg equilibrium_tol (:95-104)
  solvers.def(
      "newton_cg",
      [](Cell& c, rarr g, SolverBase& so, Real nt, Real eqt, Dim_t verb, Uint nmax) {
        return solve_driver(false, c, g, so, nt, eqt, verb, nmax);
      },
      "cell"_a, "ΔF₀"_a, "solver"_a, "newton_tol"_a, "equil_tol"_a, "verbose"_a = 0,
      py::kw_only(), "newton_maxiter"_a = 0);
  solvers.def(
      "newton_cg",
      [](Cell& c, py::list g, SolverBase& so, Real nt, Real eqt, Dim_t verb, Uint nmax) {
        return solve_driver(false, c, g, so, nt, eqt, verb, nmax);
      },
      "cell"_a, "ΔF₀"_a, "solver"_a, "newton_tol"_a, "equilibrium_tol"_a, "verbose"_a = 0,
      py::kw_only(), "newton_maxiter"_a = 0);
  solvers.def(
      "de_geus",
      [](Cell& c, rarr g, SolverBase& so, Real nt, Real eqt, Dim_t verb, Uint nmax) {
        return solve_driver(true, c, g, so, nt, eqt, verb, nmax);
      },
      "cell"_a, "ΔF₀"_a, "solver"_a, "newton_tol"_a, "equilibrium_tol"_a, "verbose"_a = 0,
      py::kw_only(), "newton_maxiter"_a = 0);
  solvers.def(
      "de_geus",
      [](Cell& c, py::list g, SolverBase& so, Real nt, Real eqt, Dim_t verb, Uint nmax) {
        return solve_driver(true, c, g, so, nt, eqt, verb, nmax);
      },
      "cell"_a, "ΔF₀"_a, "solver"_a, "newton_tol"_a, "equilibrium_tol"_a, "verbose"_a = 0,
      py::kw_only(), "newton_maxiter"_a = 0);

  // ---- fft engines and projections (bind_py_fftengine.cc, bind_py_projections.cc) ----
  auto fft = mod.def_submodule("fft");
  fft.doc() = "bindings for µSpectre's fft engines";
  bind_engine<2>(fft, "HIP_2d");
  bind_engine<3>(fft, "HIP_3d");
  fft.attr("FFTW_2d") = fft.attr("HIP_2d");
  fft.attr("FFTW_3d") = fft.attr("HIP_3d");
  bind_projection<ProjectionSmallStrain<2, 2>, 2>(fft, "ProjectionSmallStrain");
  bind_projection<ProjectionSmallStrain<3, 3>, 3>(fft, "ProjectionSmallStrain");
  bind_projection<ProjectionFiniteStrain<2, 2>, 2>(fft, "ProjectionFiniteStrain");
  bind_projection<ProjectionFiniteStrain<3, 3>, 3>(fft, "ProjectionFiniteStrain");
  bind_projection<ProjectionFiniteStrainFast<2, 2>, 2>(fft, "ProjectionFiniteStrainFast");
  bind_projection<ProjectionFiniteStrainFast<3, 3>, 3>(fft, "ProjectionFiniteStrainFast");
  // the projection a cell returns from get_projection (bind_py_cell.cc:197-198)
  {
    py::class_<CellProjection<2>> c2(fft, "CellProjection_2d");
    bind_projection_methods<CellProjection<2>, 2>(c2);
    py::class_<CellProjection<3>> c3(fft, "CellProjection_3d");
    bind_projection_methods<CellProjection<3>, 3>(c3);
  }
}
