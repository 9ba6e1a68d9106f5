// Eigen's iterative solvers as bound by the reference (src/solver/
// solver_eigen.{hh,cc}): ConjugateGradient, BiCGSTAB, GMRES, DGMRES and
// MINRES of Eigen 3.3 (the version the reference's build requires,
// CMakeLists.txt:103), each with the identity preconditioner.  The matrix
// is the cell's projected directional stiffness, applied on the device;
// vectors stay on the device and the host keeps the scalar recurrences.
#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <utility>
#include <vector>

#include "muspectre_host.hh"

namespace muSpectre {

SolverEigenBase::SolverEigenBase(Cell& c, Real t, Uint m, bool v)
    : SolverBase(c, t, m, v), x_k(c.get_nb_dof()) {}

void SolverEigenBase::apply(const DeviceVector& in, DeviceVector& out) {
  cell.evaluate_projected_directional_stiffness(in, out);
}

Real SolverEigenBase::dot(const DeviceVector& a, const DeviceVector& b) const {
  double s = 0;
  check(mugpu_dot(a.data(), b.data(), a.size(), &s));
  return s;
}

void SolverEigenBase::axpby(Real a, const DeviceVector& x, Real b, DeviceVector& y) const {
  check(mugpu_axpby(a, x.data(), b, y.data(), x.size()));
}

void SolverEigenBase::scale_copy(Real a, const DeviceVector& x, DeviceVector& y) const {
  y.resize(x.size());
  check(mugpu_scale_copy(a, x.data(), y.data(), x.size()));
}

// SolverEigen::solve (solver_eigen.cc:61-86): the solver starts from x = 0
// (IterativeSolverBase::solve); info() is NumericalIssue on a breakdown,
// Success when the relative residual is within tol, NoConvergence otherwise
DeviceVector& SolverEigenBase::solve(const DeviceVector& rhs) {
  converged = false;
  Uint its = 0;
  Real err = 0;
  x_k.resize(rhs.size());
  const bool ok = run(rhs, x_k, its, err);
  counter += its;
  if (!ok || !(err <= tol)) {
    std::stringstream e;
    e << get_name() << " has not converged,"
      << " After " << its << " steps, the solver "
      << " FAILED with  |r|/|b| = " << std::setw(15) << err << ", cg_tol = " << tol << std::endl;
    throw ConvergenceError(e.str());
  }
  if (verbose)
    std::cout << " After " << its << " " << get_name() << " steps, |r|/|b| = " << std::setw(15)
              << err << ", cg_tol = " << tol << std::endl;
  converged = true;
  return x_k;
}

// Eigen::internal::conjugate_gradient (Eigen 3.3, IterativeLinearSolvers/
// ConjugateGradient.h), identity preconditioner: stops when |r|^2 <
// max(tol^2 |b|^2, smallest normal double); the iteration that converges is
// not counted
bool SolverCGEigen::run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) {
  const int64_t n = b.size();
  check(mugpu_memset_zero(x.data(), sizeof(double) * n));
  const Real rhs2 = dot(b, b);
  if (rhs2 == 0) {
    iters = 0;
    error = 0;
    return true;
  }
  const Real threshold = std::max(tol * tol * rhs2, (std::numeric_limits<Real>::min)());
  DeviceVector r(n), p(n), t(n);
  scale_copy(1.0, b, r);  // r = b - A 0
  Real res2 = dot(r, r);
  if (res2 < threshold) {
    iters = 0;
    error = std::sqrt(res2 / rhs2);
    return true;
  }
  scale_copy(1.0, r, p);
  Real abs_new = res2;
  Uint i = 0;
  while (i < maxiter) {
    apply(p, t);
    const Real alpha = abs_new / dot(p, t);
    axpby(alpha, p, 1.0, x);
    axpby(-alpha, t, 1.0, r);
    res2 = dot(r, r);
    if (res2 < threshold) break;
    const Real abs_old = abs_new;
    abs_new = res2;
    axpby(1.0, r, abs_new / abs_old, p);
    ++i;
  }
  error = std::sqrt(res2 / rhs2);
  iters = i;
  return true;
}

// Eigen::internal::bicgstab (Eigen 3.3, IterativeLinearSolvers/BiCGSTAB.h),
// identity preconditioner, with its restart when r becomes orthogonal to the
// shadow residual r0
bool SolverBiCGSTABEigen::run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) {
  const int64_t n = b.size();
  check(mugpu_memset_zero(x.data(), sizeof(double) * n));
  const Real rhs2 = dot(b, b);
  if (rhs2 == 0) {
    iters = 0;
    error = 0;
    return true;
  }
  DeviceVector r(n), r0(n), v(n), p(n), s(n), t(n);
  scale_copy(1.0, b, r);  // r = b - A 0
  scale_copy(1.0, r, r0);
  Real r0_2 = dot(r0, r0);
  check(mugpu_memset_zero(v.data(), sizeof(double) * n));
  check(mugpu_memset_zero(p.data(), sizeof(double) * n));
  Real rho = 1, alpha = 1, w = 1;
  const Real tol2 = tol * tol * rhs2;
  const Real eps = std::numeric_limits<Real>::epsilon();
  const Real eps2 = eps * eps;
  Uint i = 0;
  int restarts = 0;
  Real r2 = dot(r, r);
  while (r2 > tol2 && i < maxiter) {
    const Real rho_old = rho;
    rho = dot(r0, r);
    if (std::abs(rho) < eps2 * r0_2) {
      // r became too orthogonal to r0: restart with r0 = b - A x
      apply(x, t);
      scale_copy(1.0, b, r);
      axpby(-1.0, t, 1.0, r);
      scale_copy(1.0, r, r0);
      rho = r0_2 = dot(r, r);
      if (restarts++ == 0) i = 0;
    }
    const Real beta = (rho / rho_old) * (alpha / w);
    // p = r + beta (p - w v)
    axpby(-w, v, 1.0, p);
    axpby(1.0, r, beta, p);
    apply(p, v);  // y = p
    alpha = rho / dot(r0, v);
    scale_copy(1.0, r, s);
    axpby(-alpha, v, 1.0, s);
    apply(s, t);  // z = s
    const Real tt = dot(t, t);
    w = tt > 0 ? dot(t, s) / tt : 0.0;
    axpby(alpha, p, 1.0, x);
    axpby(w, s, 1.0, x);
    scale_copy(1.0, s, r);
    axpby(-w, t, 1.0, r);
    r2 = dot(r, r);
    ++i;
  }
  error = std::sqrt(r2 / rhs2);
  iters = i;
  return std::isfinite(error);
}

// Restarted GMRES(restart) as Eigen::internal::gmres (Eigen 3.3,
// unsupported/Eigen/src/IterativeSolvers/GMRES.h): each inner step counts
// one iteration, the residual estimate |g_k| / |r_0| (r_0 = b for x = 0)
// is compared with tol, and at a restart x is updated and r = b - A x
// recomputed.  Eigen builds the Arnoldi basis with Householder reflections;
// here it is built with modified Gram-Schmidt on the device and the
// Hessenberg least-squares problem is reduced with Givens rotations on the
// host: the same iterates in exact arithmetic.
//
// DGMRES (Eigen 3.3, unsupported/Eigen/src/IterativeSolvers/DGMRES.h) is
// bound by the reference with its defaults (solver_eigen.hh:93-100; nothing
// in the reference calls setEigenv), and by default it deflates no
// eigenvalues: DGMRES::dgmresCycle then computes no deflation data and each
// cycle is the restarted GMRES(30) cycle below, which SolverDGMRESEigen
// runs.  One difference is deliberate: Eigen's DGMRES replaces a zero
// initial guess by the preconditioner applied to the right-hand side (x0 = b
// for the identity preconditioner).  For the Newton systems of a cell, b is a
// stress-sized residual and x a strain increment; with moduli in Pa, x0 = b
// starts from a residual about |K| |b| and a relative tolerance of 1e-8 is
// then below what FP64 can resolve.  Both solvers start from x0 = 0.
bool SolverGMRESEigen::run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) {
  const int64_t n = b.size();
  check(mugpu_memset_zero(x.data(), sizeof(double) * n));
  iters = 0;
  const Real r0_norm = std::sqrt(dot(b, b));
  if (r0_norm == 0) {
    error = 0;
    return true;
  }
  const Uint m = std::max<Uint>(1, restart);
  std::vector<DeviceVector> V;
  V.reserve(m + 1);
  for (Uint i = 0; i <= m; ++i) V.emplace_back(n);
  DeviceVector r(n), w(n);
  scale_copy(1.0, b, r);
  Real beta = r0_norm;
  error = 1.0;
  std::vector<Real> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  auto h = [&](Uint i, Uint j) -> Real& { return H[(size_t)i * m + j]; };
  while (true) {
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    scale_copy(1.0 / beta, r, V[0]);
    Uint k = 0;
    bool stop = false;
    while (k < m) {
      ++iters;
      apply(V[k], w);
      for (Uint i = 0; i <= k; ++i) {
        h(i, k) = dot(w, V[i]);
        axpby(-h(i, k), V[i], 1.0, w);
      }
      const Real hk = std::sqrt(dot(w, w));
      h(k + 1, k) = hk;
      if (hk != 0) scale_copy(1.0 / hk, w, V[k + 1]);
      for (Uint i = 0; i < k; ++i) {
        const Real a = h(i, k), c = h(i + 1, k);
        h(i, k) = cs[i] * a + sn[i] * c;
        h(i + 1, k) = -sn[i] * a + cs[i] * c;
      }
      const Real a = h(k, k), c = h(k + 1, k);
      const Real rr = std::hypot(a, c);
      cs[k] = rr != 0 ? a / rr : 1.0;
      sn[k] = rr != 0 ? c / rr : 0.0;
      h(k, k) = rr;
      h(k + 1, k) = 0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      ++k;
      error = std::abs(g[k]) / r0_norm;
      stop = error < tol || iters >= maxiter || hk == 0 || !std::isfinite(error);
      if (stop) break;
    }
    // x += V y with H y = g (upper triangular, k x k)
    for (int i = (int)k - 1; i >= 0; --i) {
      Real s = g[i];
      for (Uint j = i + 1; j < k; ++j) s -= h(i, j) * y[j];
      y[i] = h(i, i) != 0 ? s / h(i, i) : 0.0;
    }
    for (Uint i = 0; i < k; ++i) axpby(y[i], V[i], 1.0, x);
    if (stop) return std::isfinite(error);
    // restart: r = b - A x
    apply(x, w);
    scale_copy(1.0, b, r);
    axpby(-1.0, w, 1.0, r);
    beta = std::sqrt(dot(r, r));
    if (beta == 0) {
      error = 0;
      return true;
    }
  }
}

// Eigen::internal::minres (Eigen 3.3, unsupported/Eigen/src/
// IterativeSolvers/MINRES.h), identity preconditioner: Lanczos with Givens
// rotations; the estimated residual |r|^2 *= s^2 per iteration is compared
// with tol^2 |b|^2, and the iteration that converges is not counted
bool SolverMINRESEigen::run(const DeviceVector& b, DeviceVector& x, Uint& iters, Real& error) {
  const int64_t n = b.size();
  check(mugpu_memset_zero(x.data(), sizeof(double) * n));
  const Real rhs2 = dot(b, b);
  if (rhs2 == 0) {
    iters = 0;
    error = 0;
    return true;
  }
  const Real threshold2 = tol * tol * rhs2;
  DeviceVector v_old(n), v(n), v_new(n), w(n), p_oold(n), p_old(n), p(n), t(n);
  check(mugpu_memset_zero(v.data(), sizeof(double) * n));
  check(mugpu_memset_zero(p_old.data(), sizeof(double) * n));
  check(mugpu_memset_zero(p.data(), sizeof(double) * n));
  scale_copy(1.0, b, v_new);  // b - A 0
  Real res2 = dot(v_new, v_new);
  Real beta_new = std::sqrt(res2);
  const Real beta_one = beta_new;
  scale_copy(1.0 / beta_new, v_new, v_new);
  Real c = 1, c_old = 1, s = 0, s_old = 0, eta = 1;
  iters = 0;
  while (iters < maxiter) {
    const Real beta = beta_new;
    std::swap(v_old, v);
    std::swap(v, v_new);  // v = the normalised Lanczos vector, w = v
    scale_copy(1.0, v, w);
    apply(w, v_new);
    axpby(-beta, v_old, 1.0, v_new);
    const Real alpha = dot(v_new, w);
    axpby(-alpha, v, 1.0, v_new);
    beta_new = std::sqrt(dot(v_new, v_new));
    if (beta_new != 0) scale_copy(1.0 / beta_new, v_new, v_new);
    const Real r2 = s * alpha + c * c_old * beta;
    const Real r3 = s_old * beta;
    const Real r1_hat = c * alpha - c_old * s * beta;
    const Real r1 = std::sqrt(r1_hat * r1_hat + beta_new * beta_new);
    c_old = c;
    s_old = s;
    c = r1_hat / r1;
    s = beta_new / r1;
    // p = (w - r2 p_old - r3 p_oold) / r1
    std::swap(p_oold, p_old);
    std::swap(p_old, p);
    scale_copy(1.0 / r1, w, p);
    axpby(-r2 / r1, p_old, 1.0, p);
    axpby(-r3 / r1, p_oold, 1.0, p);
    axpby(beta_one * c * eta, p, 1.0, x);
    res2 *= s * s;
    if (res2 < threshold2 || !std::isfinite(res2)) break;
    eta = -s * eta;
    ++iters;
  }
  error = std::sqrt(res2 / rhs2);
  return std::isfinite(error);
}

}  // namespace muSpectre
