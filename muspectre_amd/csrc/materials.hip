// Per-voxel constitutive kernels (FP64): one launch per material over its
// own pixels in insertion order, as MaterialMuSpectre::compute_stresses_worker
// iterates them (src/materials/material_muSpectre_base.hh:292-401).
//
//   LE1  src/materials/material_linear_elastic1.hh:166-194, .cc:40-48
//   LE2  material_linear_elastic2.hh:178-201 (LE1 on strain - eigenstrain)
//   LE3  material_linear_elastic3.hh:181-198 (per-pixel C from E, nu)
//   LE4  material_linear_elastic4.hh:189-210 (per-pixel lambda, mu)
//   HEP1 material_hyper_elasto_plastic1.hh:227-309, .cc:40-84
//   finite strain: Green-Lagrange strain and PK2 -> PK1 (materials_toolbox.hh:
//   139-148,313-350); Kirchhoff -> PK1 as coded (:374-410)
//
// The tangent is written in the representation the cell chose: ISO
// (lambda, mu), HEP (dim^3, j-independent Kirchhoff->PK1 tangent), FULL
// (the reference's nc x nc matrix) or SYM (its upper triangle), in the
// voxel-blocked layout of tan_off().
#include <cmath>

#include "mugpu_internal.hpp"

namespace mugpu {

namespace {

template <int D>
__device__ __forceinline__ void hooke_stress(double lam, double mu, const double* E, double* S) {
  double tr = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) tr += E[i * (D + 1)];
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) S[i + D * j] = (i == j ? tr * lam : 0.0) + 2 * mu * E[i + D * j];
}

// C : E for C = lambda Itrac + 2 mu Isymm
template <int D>
__device__ __forceinline__ void iso_contract(double lam, double mu, const double* E, double* S) {
  double tr = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) tr += E[i * (D + 1)];
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i)
      S[i + D * j] = (i == j ? lam * tr : 0.0) + mu * (E[i + D * j] + E[j + D * i]);
}

template <int D>
__device__ __forceinline__ double iso_C(double lam, double mu, int i, int j, int k, int l) {
  const double itrac = (i == j && k == l) ? 1.0 : 0.0;
  const double isym = 0.5 * (((i == k && j == l) ? 1.0 : 0.0) + ((i == l && j == k) ? 1.0 : 0.0));
  return lam * itrac + 2 * mu * isym;
}

template <int D>
__device__ __forceinline__ void green_lagrange(const double* F, double* E) {
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += F[k + D * i] * F[k + D * j];
      E[i + D * j] = 0.5 * (s - (i == j ? 1.0 : 0.0));
    }
}

template <int D>
__device__ __forceinline__ void inverse(const double* A, double* inv) {
  if (D == 2) {
    const double det = A[0] * A[3] - A[2] * A[1];
    inv[0] = A[3] / det;
    inv[3] = A[0] / det;
    inv[2] = -A[2] / det;
    inv[1] = -A[1] / det;
  } else {
#define a(i, j) A[(i) + 3 * (j)]
    const double c00 = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
    const double c01 = a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2);
    const double c02 = a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0);
    const double det = a(0, 0) * c00 + a(0, 1) * c01 + a(0, 2) * c02;
    const double id = 1.0 / det;
    inv[0] = c00 * id;
    inv[1] = c01 * id;
    inv[2] = c02 * id;
    inv[3] = (a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2)) * id;
    inv[4] = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) * id;
    inv[5] = (a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1)) * id;
    inv[6] = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) * id;
    inv[7] = (a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)) * id;
    inv[8] = (a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)) * id;
#undef a
  }
}

template <int D>
__device__ __forceinline__ void matmul(const double* A, const double* B, double* C) {
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += A[i + D * k] * B[k + D * j];
      C[i + D * j] = s;
    }
}

// eigenvalues of a symmetric matrix, ascending (3x3: trigonometric solution
// of the characteristic cubic, O. K. Smith, CACM 4(4), 1961)
template <int D>
__device__ void eig_sym(const double* A, double* ev) {
  if (D == 2) {
    const double a = A[0], b = 0.5 * (A[2] + A[1]), d = A[3];
    const double m = 0.5 * (a + d), h = 0.5 * (a - d);
    const double r = sqrt(h * h + b * b);
    ev[0] = m - r;
    ev[1] = m + r;
    return;
  }
  const double a00 = A[0], a11 = A[4], a22 = A[8];
  const double a01 = A[3], a02 = A[6], a12 = A[7];
  const double m = (a00 + a11 + a22) / 3.0;
  const double b00 = a00 - m, b11 = a11 - m, b22 = a22 - m;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * (a01 * a01 + a02 * a02 + a12 * a12);
  const double p = p2 / 6.0;
  if (p <= 0.0) {
    ev[0] = ev[1] = ev[2] = m;
    return;
  }
  const double detB = b00 * (b11 * b22 - a12 * a12) - a01 * (a01 * b22 - a12 * a02) +
                      a02 * (a01 * a12 - b11 * a02);
  const double sp = sqrt(p);
  double r = 0.5 * detB / (p * sp);
  r = r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r);
  const double phi = acos(r) / 3.0;
  double hi = m + 2.0 * sp * cos(phi);
  double lo = m + 2.0 * sp * cos(phi + 2.0 * M_PI / 3.0);
  double mid = 3.0 * m - hi - lo;
  double e0 = lo, e1 = mid, e2 = hi, t;
  if (e1 < e0) { t = e0; e0 = e1; e1 = t; }
  if (e2 < e1) { t = e1; e1 = e2; e2 = t; }
  if (e1 < e0) { t = e0; e0 = e1; e1 = t; }
  ev[0] = e0;
  ev[1] = e1;
  ev[2] = e2;
}

__device__ __forceinline__ double log_dd1(double a, double b) {
  const double h = b - a;
  return h == 0.0 ? 1.0 / a : log1p(h / a) / h;
}
__device__ __forceinline__ double exp_dd1(double a, double b) {
  const double h = b - a;
  return h == 0.0 ? exp(a) : exp(a) * expm1(h) / h;
}

// f(T) = d0 I + d1 (T - l0 I) + d2 (T - l0 I)(T - l1 I): the Newton form of
// Sylvester's interpolation over the eigenvalues (same function in exact
// arithmetic as src/common/eigen_tools.hh:221-341, finite at repeated
// eigenvalues)
template <int D>
__device__ void newton_poly(const double* T_, const double* l, double d0, double d1, double d2,
                            double* out) {
  double A[D * D], B[D * D];
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) {
      A[i + D * j] = T_[i + D * j] - (i == j ? l[0] : 0.0);
      B[i + D * j] = T_[i + D * j] - (i == j ? l[1] : 0.0);
    }
  double AB[D * D];
  matmul<D>(A, B, AB);
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i)
      out[i + D * j] = (i == j ? d0 : 0.0) + d1 * A[i + D * j] + (D == 3 ? d2 * AB[i + D * j] : 0.0);
}

template <int D>
__device__ void logm(const double* T_, double* out) {
  double l[3];
  eig_sym<D>(T_, l);
  double d2 = 0.0;
  if (D == 3) {
    const double h = l[2] - l[0];
    d2 = h == 0.0 ? -0.5 / (l[0] * l[0]) : (log_dd1(l[1], l[2]) - log_dd1(l[0], l[1])) / h;
  }
  newton_poly<D>(T_, l, log(l[0]), log_dd1(l[0], l[1]), d2, out);
}

template <int D>
__device__ void expm(const double* T_, double* out) {
  double l[3];
  eig_sym<D>(T_, l);
  double d2 = 0.0;
  if (D == 3) {
    const double h = l[2] - l[0];
    d2 = h == 0.0 ? 0.5 * exp(l[0]) : (exp_dd1(l[1], l[2]) - exp_dd1(l[0], l[1])) / h;
  }
  newton_poly<D>(T_, l, exp(l[0]), exp_dd1(l[0], l[1]), d2, out);
}

// cyclic Jacobi eigen-decomposition of a symmetric 3x3 (Golub & Van Loan,
// sec. 8.5); columns of V are the eigenvectors
__device__ void jacobi3(const double* A, double* ev, double* V) {
  double a[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      a[i][j] = 0.5 * (A[i + 3 * j] + A[j + 3 * i]);
      V[i + 3 * j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    if (off == 0.0) break;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2, k = 3 - p - q;
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double app = a[p][p], aqq = a[q][q];
      if (sweep > 3 && fabs(app) + 100.0 * fabs(apq) == fabs(app) &&
          fabs(aqq) + 100.0 * fabs(apq) == fabs(aqq)) {
        a[p][q] = a[q][p] = 0.0;
        continue;
      }
      const double theta = (aqq - app) / (2.0 * apq);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      const double akp = a[k][p], akq = a[k][q];
      a[p][p] = app - t * apq;
      a[q][q] = aqq + t * apq;
      a[p][q] = a[q][p] = 0.0;
      a[k][p] = a[p][k] = c * akp - s * akq;
      a[k][q] = a[q][k] = s * akp + c * akq;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double vp = V[i + 3 * p], vq = V[i + 3 * q];
        V[i + 3 * p] = c * vp - s * vq;
        V[i + 3 * q] = s * vp + c * vq;
      }
    }
  }
  ev[0] = a[0][0];
  ev[1] = a[1][1];
  ev[2] = a[2][2];
}

struct MatArgs {
  const int* pix;
  int64_t n;
  double young, poisson, lambda, mu, bulk, tau_y0, H;
  const double* params;
  const double* F;
  double* P;
  double* tan;
  int tan_kind;
  int need_tangent;
  double *Fp_cur, *be_cur, *ep_cur;
  const double *Fp_old, *be_old, *ep_old;
};

// one entry K(r, c) of an nc x nc tangent in the FULL or SYM representation
template <int D>
__device__ __forceinline__ void put_K(const MatArgs& a, int64_t v, int r, int c, double val) {
  constexpr int NC = D * D;
  if (a.tan_kind == TAN_SYM) {
    if (r <= c) a.tan[tan_off(v, sym_index(r, c, NC), NC * (NC + 1) / 2)] = val;
  } else {
    a.tan[tan_off(v, r + NC * c, NC * NC)] = val;
  }
}

// consistent tangent dP/dF of the HEP1 update (no reference counterpart; the
// reference's coded tangent is not the derivative of P):
//   K(i,m,j,o) = sum_k E(i,k,j,o) Finv_mk - P_io Finv_mj,
//   E(.,.,j,o) = C : M_jo,  M_jo = sum_q W_oq [L(., (j,q)) + L(., (q,j))]
// with W = F^-1 be*, L the Daleckii-Krein derivative of ln at be* and C the
// algorithmic Kirchhoff modulus, which is c1 I(x)I + c2 Isymm + c3 N(x)N.
// In the eigenbasis V of be* (g the divided differences of ln):
//   M_jo = V [g o (V^T X V)] V^T,  X = e_j (x) W_o + W_o (x) e_j,
// so each column (j, o) of K costs a few 3x3 products and no 81-entry array.
template <int D>
__device__ void hep1_consistent_tangent(const MatArgs& a, int64_t v, const double* Fi,
                                        const double* bes, double c1, double c2, double c3,
                                        const double* Ns, const double* P) {
  constexpr int NC = D * D;
  double W[NC], ev[3], V[9];
  matmul<D>(Fi, bes, W);
  if (D == 3) {
    jacobi3(bes, ev, V);
  } else {
    const double a0 = bes[0], b0 = 0.5 * (bes[2] + bes[1]), c0 = bes[3];
    const double th = 0.5 * atan2(2 * b0, a0 - c0);
    const double cs = cos(th), sn = sin(th);
    V[0] = cs;
    V[1] = sn;
    V[2] = -sn;
    V[3] = cs;
    ev[0] = cs * cs * a0 + 2 * cs * sn * b0 + sn * sn * c0;
    ev[1] = sn * sn * a0 - 2 * cs * sn * b0 + cs * cs * c0;
  }
  double g[NC];
#pragma unroll
  for (int m = 0; m < D; ++m)
#pragma unroll
    for (int n = 0; n < D; ++n) {
      const double h = ev[n] - ev[m];
      g[m + D * n] = h == 0.0 ? 1.0 / ev[m] : log1p(h / ev[m]) / h;
    }
  // w(o, m) = sum_q W(o, q) V(q, m)
  double w[NC];
#pragma unroll
  for (int m = 0; m < D; ++m)
#pragma unroll
    for (int o = 0; o < D; ++o) {
      double s = 0;
#pragma unroll
      for (int q = 0; q < D; ++q) s += W[o + D * q] * V[q + D * m];
      w[o + D * m] = s;
    }
  for (int o = 0; o < D; ++o)
    for (int j = 0; j < D; ++j) {
      double Z[NC], M[NC], E[NC];
#pragma unroll
      for (int n = 0; n < D; ++n)
#pragma unroll
        for (int m = 0; m < D; ++m)
          Z[m + D * n] =
              g[m + D * n] * (V[j + D * m] * w[o + D * n] + w[o + D * m] * V[j + D * n]);
#pragma unroll
      for (int bb = 0; bb < D; ++bb)
#pragma unroll
        for (int aa = 0; aa < D; ++aa) {
          double s = 0;
#pragma unroll
          for (int n = 0; n < D; ++n) {
            double t = 0;
#pragma unroll
            for (int m = 0; m < D; ++m) t += V[aa + D * m] * Z[m + D * n];
            s += t * V[bb + D * n];
          }
          M[aa + D * bb] = s;
        }
      double tr = 0, nm = 0;
#pragma unroll
      for (int aa = 0; aa < D; ++aa) tr += M[aa * (D + 1)];
#pragma unroll
      for (int q = 0; q < NC; ++q) nm += Ns[q] * M[q];
#pragma unroll
      for (int k = 0; k < D; ++k)
#pragma unroll
        for (int r = 0; r < D; ++r)
          E[r + D * k] = (r == k ? c1 * tr : 0.0) + c2 * 0.5 * (M[r + D * k] + M[k + D * r]) +
                         c3 * Ns[r + D * k] * nm;
#pragma unroll
      for (int m = 0; m < D; ++m)
#pragma unroll
        for (int i = 0; i < D; ++i) {
          double s = 0;
#pragma unroll
          for (int k = 0; k < D; ++k) s += E[i + D * k] * Fi[m + D * k];
          put_K<D>(a, v, i + D * m, j + D * o, s - P[i + D * o] * Fi[m + D * j]);
        }
    }
}

// write an isotropic tangent (lambda, mu) in the cell's representation
template <int D>
__device__ void store_iso(const MatArgs& a, int64_t v, double lam, double mu) {
  if (a.tan_kind == TAN_ISO) {
    a.tan[tan_off(v, 0, 2)] = lam;
    a.tan[tan_off(v, 1, 2)] = mu;
  } else {
#pragma unroll
    for (int l = 0; l < D; ++l)
#pragma unroll
      for (int k = 0; k < D; ++k)
#pragma unroll
        for (int j = 0; j < D; ++j)
#pragma unroll
          for (int i = 0; i < D; ++i) put_K<D>(a, v, i + D * j, k + D * l, iso_C<D>(lam, mu, i, j, k, l));
  }
}

// S = C : E for a general C, C_ijkl at (i + D j) + D^2 (k + D l)
// (Matrices::tensmult, as material_linear_elastic_generic.hh:164-176)
template <int D>
__device__ __forceinline__ void general_contract(const double* __restrict__ C, const double* E,
                                                 double* S) {
  constexpr int NC = D * D;
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) s += C[r + NC * c] * E[c];
    S[r] = s;
  }
}

// small strain: the tangent is C itself
template <int D>
__device__ void store_general(const MatArgs& a, int64_t v, const double* __restrict__ C) {
  constexpr int NC = D * D;
  for (int c = 0; c < NC; ++c)
    for (int r = 0; r < NC; ++r) put_K<D>(a, v, r, c, C[r + NC * c]);
}

// PK2 -> PK1 with a general C (materials_toolbox.hh:332-349)
template <int D>
__device__ void pk1_from_pk2_store_general(const MatArgs& a, int64_t v, const double* F,
                                           const double* S, const double* __restrict__ C) {
  constexpr int NC = D * D;
  double* P = a.P + v * NC;
  double Pl[NC];
  matmul<D>(F, S, Pl);
#pragma unroll
  for (int q = 0; q < NC; ++q) P[q] = Pl[q];
  if (!a.need_tangent) return;
  for (int n = 0; n < D; ++n)
    for (int j = 0; j < D; ++j)
      for (int m = 0; m < D; ++m)
        for (int i = 0; i < D; ++i) {
          double k = (i == j) ? S[m + D * n] : 0.0;
          for (int r = 0; r < D; ++r)
            for (int s = 0; s < D; ++s)
              k += F[i + D * r] * C[(r + D * m) + NC * (n + D * s)] * F[j + D * s];
          put_K<D>(a, v, i + D * m, j + D * n, k);
        }
}

// PK2 -> PK1 with C = lambda Itrac + 2 mu Isymm (materials_toolbox.hh:332-349)
template <int D>
__device__ void pk1_from_pk2_store(const MatArgs& a, int64_t v, const double* F, const double* S,
                                   double lam, double mu) {
  constexpr int NC = D * D;
  double* P = a.P + v * NC;
  double Pl[NC];
  matmul<D>(F, S, Pl);
#pragma unroll
  for (int q = 0; q < NC; ++q) P[q] = Pl[q];
  if (!a.need_tangent) return;
  for (int n = 0; n < D; ++n)
    for (int j = 0; j < D; ++j)
      for (int m = 0; m < D; ++m)
        for (int i = 0; i < D; ++i) {
          double k = (i == j) ? S[m + D * n] : 0.0;
          for (int r = 0; r < D; ++r)
            for (int s = 0; s < D; ++s) k += F[i + D * r] * iso_C<D>(lam, mu, r, m, n, s) * F[j + D * s];
          put_K<D>(a, v, i + D * m, j + D * n, k);
        }
}

template <int D, int FAM, int FORM>
__global__ __launch_bounds__(256) void k_linear_elastic(MatArgs a) {
  constexpr int NC = D * D;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const int64_t v = a.pix[q];
  double F[NC], E[NC], S[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) F[c] = a.F[v * NC + c];
  double lam = a.lambda, mu = a.mu;
  if (FAM == MUGPU_LE3 || FAM == MUGPU_LE4) {
    lam = a.params[2 * q];
    mu = a.params[2 * q + 1];
  }
  if (FORM == MUGPU_SMALL_STRAIN) {
#pragma unroll
    for (int c = 0; c < NC; ++c) E[c] = F[c];
  } else {
    green_lagrange<D>(F, E);
  }
  if (FAM == MUGPU_LE2) {
#pragma unroll
    for (int c = 0; c < NC; ++c) E[c] -= a.params[q * NC + c];
  }
  if (FAM == MUGPU_LEG) {
    // one C for the whole material (a.params)
    general_contract<D>(a.params, E, S);
    if (FORM == MUGPU_SMALL_STRAIN) {
#pragma unroll
      for (int c = 0; c < NC; ++c) a.P[v * NC + c] = S[c];
      if (a.need_tangent) store_general<D>(a, v, a.params);
    } else {
      pk1_from_pk2_store_general<D>(a, v, F, S, a.params);
    }
    return;
  }
  if (FAM == MUGPU_LE1 || FAM == MUGPU_LE2) hooke_stress<D>(lam, mu, E, S);
  else iso_contract<D>(lam, mu, E, S);
  if (FORM == MUGPU_SMALL_STRAIN) {
#pragma unroll
    for (int c = 0; c < NC; ++c) a.P[v * NC + c] = S[c];
    if (a.need_tangent) store_iso<D>(a, v, lam, mu);
  } else {
    pk1_from_pk2_store<D>(a, v, F, S, lam, mu);
  }
}

template <int D, bool CT>
__global__ __launch_bounds__(256) void k_hep1(MatArgs a) {
  constexpr int NC = D * D;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const int64_t v = a.pix[q];
  const double lam = a.lambda, mu = a.mu;
  double F[NC], Fpo[NC], beo[NC], Fpi[NC], f[NC], tmp[NC], bes[NC], lnb[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    F[c] = a.F[v * NC + c];
    Fpo[c] = a.Fp_old[q * NC + c];
    beo[c] = a.be_old[q * NC + c];
  }
  const double eps_old = a.ep_old[q];
  // trial state: f = F Fprev^-1, be* = f be f^T, ln be*
  inverse<D>(Fpo, Fpi);
  matmul<D>(F, Fpi, f);
  matmul<D>(f, beo, tmp);
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += tmp[i + D * k] * f[j + D * k];
      bes[i + D * j] = s;
    }
  logm<D>(bes, lnb);
  double ts[NC], td[NC], Ns[NC], tau[NC];
  hooke_stress<D>(lam, mu, lnb, ts);
#pragma unroll
  for (int c = 0; c < NC; ++c) ts[c] *= 0.5;
  double tr = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) tr += ts[i * (D + 1)];
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) td[i + D * j] = ts[i + D * j] - (i == j ? tr / D : 0.0);
  double ss = 0;
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) ss += td[i + D * j] * td[j + D * i];
  const double taueq = sqrt(3 * .5 * ss);
#pragma unroll
  for (int c = 0; c < NC; ++c) Ns[c] = 3 * .5 * td[c] / taueq;
  double phi = taueq - a.tau_y0 - a.H * eps_old;
  phi = (phi < 0.0) ? 0.0 : phi;
  const double dg = phi / (a.H + 3 * mu);
#pragma unroll
  for (int c = 0; c < NC; ++c) tau[c] = ts[c] - 2 * dg * mu * Ns[c];
  // internal variable updates, as coded (eps_p.current() += dgamma)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    a.Fp_cur[q * NC + c] = F[c];
    lnb[c] -= 2 * dg * Ns[c];
  }
  double be_new[NC];
  expm<D>(lnb, be_new);
#pragma unroll
  for (int c = 0; c < NC; ++c) a.be_cur[q * NC + c] = be_new[c];
  a.ep_cur[q] += dg;
  // P = tau F^-T
  double Fi[NC];
  inverse<D>(F, Fi);
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += tau[i + D * k] * Fi[j + D * k];
      a.P[v * NC + i + D * j] = s;
    }
  if (!a.need_tangent) return;
  const bool plastic = phi > 0;
  const double a0 = plastic ? dg * mu / taueq : 0.0;
  const double a1 = mu / (a.H + 3 * mu);
  if (CT) {
    // algorithmic modulus C = c1 Itrac + c2 Isymm + c3 N (x) N: plastic
    // (K/2 - mu/3 + a0 mu, (1 - 3 a0) mu, 2 mu (a0 - a1)), elastic half the
    // Hooke tensor (lambda/2, mu, 0) (material_hyper_elasto_plastic1.hh:299-308)
    const double c1 = plastic ? a.bulk / 2. - mu / 3 + a0 * mu : 0.5 * lam;
    const double c2 = plastic ? (1 - 3 * a0) * mu : mu;
    const double c3 = plastic ? 2 * mu * (a0 - a1) : 0.0;
    double Pl[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) Pl[c] = a.P[v * NC + c];
    hep1_consistent_tangent<D>(a, v, Fi, bes, c1, c2, c3, Ns, Pl);
    return;
  }
  // Kirchhoff tangent C(r,m,n,s), then M(i,m,n) = sum_r Finv(i,r) sum_s C(r,m,n,s)
  double csum[D * D * D];  // c(r,m,n) = sum_s C(r,m,n,s) at r + D m + D^2 n
#pragma unroll
  for (int n = 0; n < D; ++n)
#pragma unroll
    for (int m = 0; m < D; ++m)
#pragma unroll
      for (int r = 0; r < D; ++r) {
        double acc = 0;
#pragma unroll
        for (int s = 0; s < D; ++s) {
          double C;
          if (plastic) {
            const double itrac = (r == m && n == s) ? 1.0 : 0.0;
            const double isym = 0.5 * (((r == n && m == s) ? 1.0 : 0.0) + ((r == s && m == n) ? 1.0 : 0.0));
            C = (a.bulk / 2. - mu / 3 + a0 * mu) * itrac + (1 - 3 * a0) * mu * isym +
                2 * mu * (a0 - a1) * Ns[r + D * m] * Ns[n + D * s];
          } else {
            C = 0.5 * iso_C<D>(lam, mu, r, m, n, s);
          }
          acc += C;
        }
        csum[r + D * m + D * D * n] = acc;
      }
  double M[D * D * D];
#pragma unroll
  for (int n = 0; n < D; ++n)
#pragma unroll
    for (int m = 0; m < D; ++m)
#pragma unroll
      for (int i = 0; i < D; ++i) {
        double acc = 0;
#pragma unroll
        for (int r = 0; r < D; ++r) acc += Fi[i + D * r] * csum[r + D * m + D * D * n];
        M[i + D * m + D * D * n] = acc;
      }
  if (a.tan_kind == TAN_HEP) {
#pragma unroll
    for (int e = 0; e < D * D * D; ++e) a.tan[tan_off(v, e, D * D * D)] = M[e];
  } else {
    for (int n = 0; n < D; ++n)
      for (int j = 0; j < D; ++j)
        for (int m = 0; m < D; ++m)
          for (int i = 0; i < D; ++i) put_K<D>(a, v, i + D * m, j + D * n, M[i + D * m + D * D * n]);
  }
}

template <int D, int FORM>
int launch_le(const MaterialDev& m, MatArgs& a, dim3 grid) {
  switch (m.family) {
    case MUGPU_LE1: hipLaunchKernelGGL((k_linear_elastic<D, MUGPU_LE1, FORM>), grid, dim3(256), 0, stream(), a); break;
    case MUGPU_LE2: hipLaunchKernelGGL((k_linear_elastic<D, MUGPU_LE2, FORM>), grid, dim3(256), 0, stream(), a); break;
    case MUGPU_LE3: hipLaunchKernelGGL((k_linear_elastic<D, MUGPU_LE3, FORM>), grid, dim3(256), 0, stream(), a); break;
    case MUGPU_LEG: hipLaunchKernelGGL((k_linear_elastic<D, MUGPU_LEG, FORM>), grid, dim3(256), 0, stream(), a); break;
    default: hipLaunchKernelGGL((k_linear_elastic<D, MUGPU_LE4, FORM>), grid, dim3(256), 0, stream(), a); break;
  }
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

// the material stage of up to kServeMatSlots serving requests of one
// resolution: per voxel the arithmetic of k_linear_elastic<D, LE4, small
// strain> over the identity pixel list, P negated into the right-hand side
// and the tangent rounded to FP32 as convert() does
struct ServeMatArgs {
  ServeMatSlot s[kServeMatSlots];
};
static_assert(sizeof(ServeMatArgs) + sizeof(int64_t) <= 4096, "kernel arguments of at most 4 KiB");
template <int D>
__global__ __launch_bounds__(256) void k_le4_serve_batch(ServeMatArgs a, int64_t n) {
  constexpr int NC = D * D;
  const ServeMatSlot& sl = a.s[blockIdx.y];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  double E[NC], S[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) E[c] = sl.F[v * NC + c];
  const double lam = sl.params[2 * v], mu = sl.params[2 * v + 1];
  iso_contract<D>(lam, mu, E, S);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    sl.P[v * NC + c] = S[c];
    sl.rhs[v * NC + c] = S[c] * -1.0;
  }
  sl.tan64[tan_off(v, 0, 2)] = lam;
  sl.tan64[tan_off(v, 1, 2)] = mu;
  sl.tan32[tan_off(v, 0, 2)] = (float)(lam * 1.0);
  sl.tan32[tan_off(v, 1, 2)] = (float)(mu * 1.0);
}

}  // namespace

int serve_material_batch(int dim, const ServeMatSlot* slots, int nslots, int64_t n) {
  if (nslots < 1 || nslots > kServeMatSlots || n < 1 || (dim != 2 && dim != 3)) {
    set_error("serve_material_batch: %d requests of %lld voxels in %dD", nslots, (long long)n, dim);
    return MUGPU_EARG;
  }
  ServeMatArgs a{};
  for (int i = 0; i < nslots; ++i) a.s[i] = slots[i];
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)nslots);
  timing_begin(MUGPU_K_MATERIAL);
  if (dim == 3) hipLaunchKernelGGL(k_le4_serve_batch<3>, grid, dim3(256), 0, stream(), a, n);
  else hipLaunchKernelGGL(k_le4_serve_batch<2>, grid, dim3(256), 0, stream(), a, n);
  timing_end(MUGPU_K_MATERIAL);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int eval_material(const MaterialDev& m, int dim, int form, const double* F, double* P,
                  int tan_kind, double* tangent, int need_tangent) {
  const int64_t n = m.n_pixels >= 0 ? m.n_pixels : (int64_t)m.h_pix.size();
  if (n == 0) return MUGPU_OK;
  MatArgs a{};
  a.pix = m.d_pix;
  a.n = n;
  a.young = m.young;
  a.poisson = m.poisson;
  a.lambda = m.lambda;
  a.mu = m.mu;
  a.bulk = m.bulk;
  a.tau_y0 = m.tau_y0;
  a.H = m.H;
  a.params = m.family == MUGPU_LEG ? m.d_C : m.d_params;
  a.F = F;
  a.P = P;
  a.tan = tangent;
  a.tan_kind = tan_kind;
  a.need_tangent = need_tangent;
  dim3 grid((unsigned)((n + 255) / 256));
  timing_begin(MUGPU_K_MATERIAL);
  struct End {
    ~End() { timing_end(MUGPU_K_MATERIAL); }
  } end_guard;
  if (m.family == MUGPU_HEP1) {
    const int cur = m.cur, old = cur ^ 1;
    a.Fp_cur = m.Fp[cur];
    a.be_cur = m.be[cur];
    a.ep_cur = m.ep[cur];
    a.Fp_old = m.Fp[old];
    a.be_old = m.be[old];
    a.ep_old = m.ep[old];
    if (dim == 3) hipLaunchKernelGGL((k_hep1<3, false>), grid, dim3(256), 0, stream(), a);
    else hipLaunchKernelGGL((k_hep1<2, false>), grid, dim3(256), 0, stream(), a);
    MUGPU_HIP(hipGetLastError());
    return MUGPU_OK;
  }
  if (m.family == MUGPU_HEP1_CT) {
    const int cur = m.cur, old = cur ^ 1;
    a.Fp_cur = m.Fp[cur];
    a.be_cur = m.be[cur];
    a.ep_cur = m.ep[cur];
    a.Fp_old = m.Fp[old];
    a.be_old = m.be[old];
    a.ep_old = m.ep[old];
    if (dim == 3) hipLaunchKernelGGL((k_hep1<3, true>), grid, dim3(256), 0, stream(), a);
    else hipLaunchKernelGGL((k_hep1<2, true>), grid, dim3(256), 0, stream(), a);
    MUGPU_HIP(hipGetLastError());
    return MUGPU_OK;
  }
  if (dim == 3)
    return form == MUGPU_SMALL_STRAIN ? launch_le<3, MUGPU_SMALL_STRAIN>(m, a, grid)
                                      : launch_le<3, MUGPU_FINITE_STRAIN>(m, a, grid);
  return form == MUGPU_SMALL_STRAIN ? launch_le<2, MUGPU_SMALL_STRAIN>(m, a, grid)
                                    : launch_le<2, MUGPU_FINITE_STRAIN>(m, a, grid);
}

}  // namespace mugpu
