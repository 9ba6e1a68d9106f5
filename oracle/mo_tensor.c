/*
 * Small symmetric-tensor algebra of the oracle (test infrastructure only).
 *
 * The reference computes logm by Sylvester's formula over the eigenvalues of
 * Eigen's SelfAdjointEigenSolver::computeDirect (src/common/eigen_tools.hh:
 * 221-322) and expm by eigen-decomposition (:329-341).  Eigen is an
 * un-vendored dependency (>= 3.3, CMakeLists.txt:103).  Here:
 *
 *  - eigenvalues of a symmetric 3x3 come from the closed-form trigonometric
 *    solution of the characteristic cubic (O. K. Smith, "Eigenvalues of a
 *    symmetric 3x3 matrix", CACM 4(4), 1961): shift by tr/3, p = |B|^2/6,
 *    q = det(B)/2, phi = acos(q/p^1.5)/3;
 *  - logm and expm are evaluated as the Newton divided-difference form of the
 *    interpolating polynomial over those eigenvalues,
 *        f(T) = f[l0] I + f[l0,l1] (T - l0 I) + f[l0,l1,l2] (T - l0 I)(T - l1 I),
 *    which equals Sylvester's Lagrange form (and the eigen-decomposition sum)
 *    in exact arithmetic, stays finite at repeated eigenvalues (confluent
 *    divided differences) and does not amplify rounding by 1/(l_i - l_j).
 *    The reference's literal Sylvester form is kept as mo_logm_sylvester for
 *    the tests that show both agree on separated spectra.
 */
#include "muspectre_oracle.h"
#include "mo_internal.h"

#include <math.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
void mo_eig_sym2(const double *A, double *ev) {
  const double a = A[0], b = 0.5 * (A[2] + A[1]), d = A[3];
  const double m = 0.5 * (a + d);
  const double h = 0.5 * (a - d);
  const double r = sqrt(h * h + b * b);
  ev[0] = m - r;
  ev[1] = m + r;
}

void mo_eig_sym3(const double *A, double *ev) {
  /* symmetric part is used (A is symmetric in all callers) */
  const double a00 = A[0], a11 = A[4], a22 = A[8];
  const double a01 = A[3], a02 = A[6], a12 = A[7];
  const double m = (a00 + a11 + a22) / 3.0;
  const double b00 = a00 - m, b11 = a11 - m, b22 = a22 - m;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 +
                    2.0 * (a01 * a01 + a02 * a02 + a12 * a12);
  const double p = p2 / 6.0;
  if (p <= 0.0) {
    ev[0] = ev[1] = ev[2] = m;
    return;
  }
  const double detB = b00 * (b11 * b22 - a12 * a12) -
                      a01 * (a01 * b22 - a12 * a02) +
                      a02 * (a01 * a12 - b11 * a02);
  const double sp = sqrt(p);
  double r = 0.5 * detB / (p * sp);
  if (r > 1.0) r = 1.0;
  if (r < -1.0) r = -1.0;
  const double phi = acos(r) / 3.0;
  const double hi = m + 2.0 * sp * cos(phi);
  const double lo = m + 2.0 * sp * cos(phi + 2.0 * M_PI / 3.0);
  const double mid = 3.0 * m - hi - lo;
  ev[0] = lo;
  ev[1] = mid;
  ev[2] = hi;
  /* keep ascending order robustly */
  if (ev[1] < ev[0]) { double t = ev[0]; ev[0] = ev[1]; ev[1] = t; }
  if (ev[2] < ev[1]) { double t = ev[1]; ev[1] = ev[2]; ev[2] = t; }
  if (ev[1] < ev[0]) { double t = ev[0]; ev[0] = ev[1]; ev[1] = t; }
}

/* eigenvectors for the literal (reference-form) expm used in tests only:
 * for each eigenvalue take the largest cross product of two rows of
 * (A - l I); orthonormalise degenerate pairs. */
void mo_eigsys_sym3(const double *A, double *ev, double *V) {
  mo_eig_sym3(A, ev);
  double used[3][3];
  int nused = 0;
  for (int k = 0; k < 3; ++k) {
    double B[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        B[i][j] = 0.5 * (A[i + 3 * j] + A[j + 3 * i]) - (i == j ? ev[k] : 0.0);
    double best[3] = {0, 0, 0}, bn = 0.0;
    for (int r0 = 0; r0 < 3; ++r0)
      for (int r1 = r0 + 1; r1 < 3; ++r1) {
        double c[3] = {B[r0][1] * B[r1][2] - B[r0][2] * B[r1][1],
                       B[r0][2] * B[r1][0] - B[r0][0] * B[r1][2],
                       B[r0][0] * B[r1][1] - B[r0][1] * B[r1][0]};
        double n = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        if (n > bn) { bn = n; memcpy(best, c, sizeof(best)); }
      }
    double v[3];
    if (bn > 1e-28) {
      double s = 1.0 / sqrt(bn);
      for (int i = 0; i < 3; ++i) v[i] = best[i] * s;
    } else {
      /* degenerate: pick a unit vector orthogonal to the used ones */
      double cand[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      double bestn = -1;
      for (int c = 0; c < 3; ++c) {
        double w[3] = {cand[c][0], cand[c][1], cand[c][2]};
        for (int u = 0; u < nused; ++u) {
          double d = w[0] * used[u][0] + w[1] * used[u][1] + w[2] * used[u][2];
          for (int i = 0; i < 3; ++i) w[i] -= d * used[u][i];
        }
        double n = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        if (n > bestn) { bestn = n; memcpy(v, w, sizeof(v)); }
      }
      double s = 1.0 / sqrt(bestn);
      for (int i = 0; i < 3; ++i) v[i] *= s;
    }
    /* orthogonalise against previously found vectors */
    for (int u = 0; u < nused; ++u) {
      double d = v[0] * used[u][0] + v[1] * used[u][1] + v[2] * used[u][2];
      for (int i = 0; i < 3; ++i) v[i] -= d * used[u][i];
    }
    double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) v[i] /= n;
    memcpy(used[nused++], v, sizeof(v));
    for (int i = 0; i < 3; ++i) V[i + 3 * k] = v[i];
  }
}

/* ---- divided differences ------------------------------------------------ */
/* log: f[a,b] = (ln b - ln a)/(b - a) = log1p((b-a)/a)/(b-a); f[a,a] = 1/a */
static double log_dd1(double a, double b) {
  const double h = b - a;
  if (h == 0.0) return 1.0 / a;
  return log1p(h / a) / h;
}
/* exp: f[a,b] = e^a expm1(b-a)/(b-a); f[a,a] = e^a */
static double exp_dd1(double a, double b) {
  const double h = b - a;
  if (h == 0.0) return exp(a);
  return exp(a) * expm1(h) / h;
}
/* second divided difference with confluent limits; f2(a) = f''(a)/2 */
static double dd2(double (*dd1)(double, double), double f2conf, double a,
                  double b, double c) {
  const double h = c - a;
  if (h == 0.0) return f2conf;
  return (dd1(b, c) - dd1(a, b)) / h;
}

/* f(T) = d0 I + d1 (T - l0 I) + d2 (T - l0 I)(T - l1 I), col-major dim x dim */
static void newton_poly(int d, const double *T, const double *l, double d0,
                        double d1, double d2, double *out) {
  double A[9], B[9], AB[9];
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      A[i + d * j] = T[i + d * j] - (i == j ? l[0] : 0.0);
      B[i + d * j] = T[i + d * j] - (i == j ? l[1] : 0.0);
    }
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double s = 0.0;
      for (int k = 0; k < d; ++k) s += A[i + d * k] * B[k + d * j];
      AB[i + d * j] = s;
    }
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i)
      out[i + d * j] = (i == j ? d0 : 0.0) + d1 * A[i + d * j] +
                       (d == 3 ? d2 * AB[i + d * j] : 0.0);
}

/* src/common/eigen_tools.hh:316-322 (same function, stable evaluation) */
void mo_logm(int d, const double *T, double *out) {
  double l[3];
  if (d == 2) {
    mo_eig_sym2(T, l);
    newton_poly(2, T, l, log(l[0]), log_dd1(l[0], l[1]), 0.0, out);
  } else {
    mo_eig_sym3(T, l);
    const double d2 = dd2(log_dd1, -0.5 / (l[0] * l[0]), l[0], l[1], l[2]);
    newton_poly(3, T, l, log(l[0]), log_dd1(l[0], l[1]), d2, out);
  }
}

/* src/common/eigen_tools.hh:329-341 (same function, stable evaluation) */
void mo_expm(int d, const double *T, double *out) {
  double l[3];
  if (d == 2) {
    mo_eig_sym2(T, l);
    newton_poly(2, T, l, exp(l[0]), exp_dd1(l[0], l[1]), 0.0, out);
  } else {
    mo_eig_sym3(T, l);
    const double d2 = dd2(exp_dd1, 0.5 * exp(l[0]), l[0], l[1], l[2]);
    newton_poly(3, T, l, exp(l[0]), exp_dd1(l[0], l[1]), d2, out);
  }
}

/* literal Sylvester form of src/common/eigen_tools.hh:221-322 (tests only):
 * sum_i ln(l_i) prod_{j != i, j = d-1 .. 0} (T - l_j I)/(l_i - l_j) */
void mo_logm_sylvester(int d, const double *T, double *out) {
  double l[3];
  if (d == 2) mo_eig_sym2(T, l); else mo_eig_sym3(T, l);
  double acc[9] = {0};
  for (int i = d - 1; i >= 0; --i) {
    double P[9];
    for (int k = 0; k < d * d; ++k) P[k] = (k % (d + 1) == 0) ? 1.0 : 0.0;
    int first = 1;
    for (int j = d - 1; j >= 0; --j) {
      if (j == i) continue;
      double M[9], R[9];
      const double s = 1.0 / (l[i] - l[j]);
      for (int c = 0; c < d; ++c)
        for (int r = 0; r < d; ++r)
          M[r + d * c] = s * (T[r + d * c] - (r == c ? l[j] : 0.0));
      if (first) {
        memcpy(P, M, sizeof(double) * d * d);
        first = 0;
      } else {
        for (int c = 0; c < d; ++c)
          for (int r = 0; r < d; ++r) {
            double v = 0;
            for (int k = 0; k < d; ++k) v += P[r + d * k] * M[k + d * c];
            R[r + d * c] = v;
          }
        memcpy(P, R, sizeof(double) * d * d);
      }
    }
    const double li = log(l[i]);
    for (int k = 0; k < d * d; ++k) acc[k] += li * P[k];
  }
  memcpy(out, acc, sizeof(double) * d * d);
}

/* literal eigen-decomposition expm (src/common/eigen_tools.hh:329-341) */
void mo_expm_eig(const double *T, double *out) {
  double l[3], V[9];
  mo_eigsys_sym3(T, l, V);
  for (int k = 0; k < 9; ++k) out[k] = 0.0;
  for (int e = 0; e < 3; ++e) {
    const double x = exp(l[e]);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) out[i + 3 * j] += x * V[i + 3 * e] * V[j + 3 * e];
  }
}

/* ---- inverse via adjugate ---------------------------------------------- */
void mo_inverse(int d, const double *A, double *inv) {
  if (d == 2) {
    const double det = A[0] * A[3] - A[2] * A[1];
    inv[0] = A[3] / det;
    inv[3] = A[0] / det;
    inv[2] = -A[2] / det;
    inv[1] = -A[1] / det;
    return;
  }
#define a(i, j) A[(i) + 3 * (j)]
  const double c00 = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
  const double c01 = a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2);
  const double c02 = a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0);
  const double det = a(0, 0) * c00 + a(0, 1) * c01 + a(0, 2) * c02;
  const double id = 1.0 / det;
  inv[0 + 3 * 0] = c00 * id;
  inv[1 + 3 * 0] = c01 * id;
  inv[2 + 3 * 0] = c02 * id;
  inv[0 + 3 * 1] = (a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2)) * id;
  inv[1 + 3 * 1] = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) * id;
  inv[2 + 3 * 1] = (a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1)) * id;
  inv[0 + 3 * 2] = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) * id;
  inv[1 + 3 * 2] = (a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)) * id;
  inv[2 + 3 * 2] = (a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)) * id;
#undef a
}

/* cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Golub & Van
 * Loan, "Matrix Computations", sec. 8.5): A = V diag(ev) V^T, columns of V
 * are the eigenvectors.  Orthonormal to rounding also for repeated
 * eigenvalues, which the consistent tangent's divided differences need. */
void mo_jacobi_sym3(const double *A, double *ev, double *V) {
  double a[3][3];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      a[i][j] = 0.5 * (A[i + 3 * j] + A[j + 3 * i]);
      V[i + 3 * j] = (i == j) ? 1.0 : 0.0;
    }
  static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    if (off == 0.0) break;
    for (int r = 0; r < 3; ++r) {
      const int p = P[r], q = Q[r];
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double app = a[p][p], aqq = a[q][q];
      /* skip rotations below the rounding of the diagonal */
      if (sweep > 3 && fabs(app) + 100.0 * fabs(apq) == fabs(app) &&
          fabs(aqq) + 100.0 * fabs(apq) == fabs(aqq)) {
        a[p][q] = a[q][p] = 0.0;
        continue;
      }
      const double theta = (aqq - app) / (2.0 * apq);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      const int k = 3 - p - q;
      const double akp = a[k][p], akq = a[k][q];
      a[p][p] = app - t * apq;
      a[q][q] = aqq + t * apq;
      a[p][q] = a[q][p] = 0.0;
      a[k][p] = a[p][k] = c * akp - s * akq;
      a[k][q] = a[q][k] = s * akp + c * akq;
      for (int i = 0; i < 3; ++i) {
        const double vp = V[i + 3 * p], vq = V[i + 3 * q];
        V[i + 3 * p] = c * vp - s * vq;
        V[i + 3 * q] = s * vp + c * vq;
      }
    }
  }
  for (int i = 0; i < 3; ++i) ev[i] = a[i][i];
}
