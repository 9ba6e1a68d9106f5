/*
 * Single-point constitutive laws of the oracle (test infrastructure only).
 * Each function restates the reference file:line it cites.
 */
#include "muspectre_oracle.h"
#include "mo_internal.h"

#include <math.h>
#include <string.h>

void mo_inverse(int d, const double *A, double *inv);
void mo_hep1_eval_kirchhoff_bes(int d, double young, double poisson,
                                double tau_y0, double H, const double *F,
                                const double *F_prev_old, const double *be_old,
                                double eps_old, double *F_prev_cur,
                                double *be_cur, double *eps_cur, double *tau,
                                double *C, double *bes_out);
void mo_logm(int d, const double *T, double *out);
void mo_expm(int d, const double *T, double *out);

/* src/materials/materials_toolbox.hh:488-525 (lambda, mu, K from E, nu) */
double mo_lambda(double E, double nu) { return E * nu / ((1 + nu) * (1 - 2 * nu)); }
double mo_mu(double E, double nu) { return E / (2 * (1 + nu)); }
double mo_bulk(double E, double nu) { return E / (3 * (1 - 2 * nu)); }

/* src/materials/materials_toolbox.hh:641-645: lambda*Itrac + 2*mu*Isymm,
 * Itrac = outer(I,I), Isymm = (outer_under(I,I)+outer_over(I,I))/2
 * (src/common/tensor_algebra.hh:155-287) */
void mo_hooke_C(int d, double lam, double mu, double *C) {
  for (int l = 0; l < d; ++l)
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) {
          const double itrac = (i == j && k == l) ? 1.0 : 0.0;
          const double isym =
              0.5 * (((i == k && j == l) ? 1.0 : 0.0) + ((i == l && j == k) ? 1.0 : 0.0));
          T4(C, i, j, k, l, d) = lam * itrac + 2 * mu * isym;
        }
}

/* src/materials/materials_toolbox.hh:660-664: tr(E)*lambda*I + 2*mu*E */
void mo_hooke_stress(int d, double lam, double mu, const double *E, double *S) {
  double tr = 0;
  for (int i = 0; i < d; ++i) tr += E[i + d * i];
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i)
      S[i + d * j] = (i == j ? tr * lam : 0.0) + 2 * mu * E[i + d * j];
}

/* src/common/tensor_algebra.hh:234-259: R_ij = sum_kl C(i,j,k,l) B_kl */
void mo_tensmult(int d, const double *C, const double *B, double *R) {
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double s = 0;
      for (int l = 0; l < d; ++l)
        for (int k = 0; k < d; ++k) s += T4(C, i, j, k, l, d) * B[k + d * l];
      R[i + d * j] = s;
    }
}

/* src/materials/materials_toolbox.hh:139-148: E = (F^T F - I)/2 */
void mo_green_lagrange(int d, const double *F, double *E) {
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double s = 0;
      for (int k = 0; k < d; ++k) s += F[k + d * i] * F[k + d * j];
      E[i + d * j] = 0.5 * (s - (i == j ? 1.0 : 0.0));
    }
}

/* src/materials/materials_toolbox.hh:313-350 */
void mo_pk1_from_pk2(int d, const double *F, const double *S, const double *C,
                     double *P, double *K) {
  for (int k = 0; k < d * d * d * d; ++k) K[k] = 0.0;
  for (int i = 0; i < d; ++i)
    for (int m = 0; m < d; ++m)
      for (int n = 0; n < d; ++n) {
        T4(K, i, m, i, n, d) += S[m + d * n];
        for (int j = 0; j < d; ++j)
          for (int r = 0; r < d; ++r)
            for (int s = 0; s < d; ++s)
              T4(K, i, m, j, n, d) += F[i + d * r] * T4(C, r, m, n, s, d) * F[j + d * s];
      }
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double v = 0;
      for (int k = 0; k < d; ++k) v += F[i + d * k] * S[k + d * j];
      P[i + d * j] = v;
    }
}

/* src/materials/materials_toolbox.hh:374-410, as coded: the tangent does
 * not depend on j, K(i,m,j,n) = sum_{r,s} Finv(i,r) C(r,m,n,s) */
void mo_pk1_from_kirchhoff(int d, const double *F, const double *tau,
                           const double *C, double *P, double *K) {
  double Fi[9];
  mo_inverse(d, F, Fi);
  for (int k = 0; k < d * d * d * d; ++k) K[k] = 0.0;
  for (int i = 0; i < d; ++i)
    for (int m = 0; m < d; ++m)
      for (int n = 0; n < d; ++n)
        for (int j = 0; j < d; ++j)
          for (int r = 0; r < d; ++r)
            for (int s = 0; s < d; ++s)
              T4(K, i, m, j, n, d) += Fi[i + d * r] * T4(C, r, m, n, s, d);
  /* P = tau * Finv^T */
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double v = 0;
      for (int k = 0; k < d; ++k) v += tau[i + d * k] * Fi[j + d * k];
      P[i + d * j] = v;
    }
}

/* LE1 (src/materials/material_linear_elastic1.hh:166-194, .cc:40-48) through
 * the formulation dispatch of material_muSpectre_base.hh:292-361 */
void mo_le1_eval(int d, int form, double young, double poisson,
                 const double *F, double *P, double *K) {
  const double lam = mo_lambda(young, poisson), mu = mo_mu(young, poisson);
  double C[81];
  mo_hooke_C(d, lam, mu, C);
  if (form == MO_SMALL_STRAIN) {
    mo_hooke_stress(d, lam, mu, F, P);
    memcpy(K, C, sizeof(double) * d * d * d * d);
  } else {
    double E[9], S[9];
    mo_green_lagrange(d, F, E);
    mo_hooke_stress(d, lam, mu, E, S);
    mo_pk1_from_pk2(d, F, S, C, P, K);
  }
}

/*
 * HEP1 stress and internals (src/materials/material_hyper_elasto_plastic1.hh:
 * 227-271) and tangent (:284-309), with C = C_T4(lambda, mu)/2 for the
 * elastic branch (.cc:62).  Writes the current() state values exactly as
 * coded: F_prev.cur = F, be.cur = expm(ln be* - 2 dgamma N*), and
 * eps_p.cur += dgamma (accumulating on the current buffer).
 */
void mo_hep1_eval_kirchhoff(int d, double young, double poisson,
                            double tau_y0, double H, const double *F,
                            const double *F_prev_old, const double *be_old,
                            double eps_old, double *F_prev_cur,
                            double *be_cur, double *eps_cur, double *tau,
                            double *C) {
  mo_hep1_eval_kirchhoff_bes(d, young, poisson, tau_y0, H, F, F_prev_old, be_old, eps_old,
                             F_prev_cur, be_cur, eps_cur, tau, C, NULL);
}

void mo_hep1_eval_kirchhoff_bes(int d, double young, double poisson,
                                double tau_y0, double H, const double *F,
                                const double *F_prev_old, const double *be_old,
                                double eps_old, double *F_prev_cur,
                                double *be_cur, double *eps_cur, double *tau,
                                double *C, double *bes_out) {
  const double lam = mo_lambda(young, poisson), mu = mo_mu(young, poisson);
  const double Kb = mo_bulk(young, poisson);
  double Fpi[9], f[9], tmp[9], bes[9], lnb[9], ts[9], td[9], Ns[9];
  mo_inverse(d, F_prev_old, Fpi);
  /* f = F * F_prev.old^-1 */
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double s = 0;
      for (int k = 0; k < d; ++k) s += F[i + d * k] * Fpi[k + d * j];
      f[i + d * j] = s;
    }
  /* be* = f * be.old * f^T */
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double s = 0;
      for (int k = 0; k < d; ++k) s += f[i + d * k] * be_old[k + d * j];
      tmp[i + d * j] = s;
    }
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double s = 0;
      for (int k = 0; k < d; ++k) s += tmp[i + d * k] * f[j + d * k];
      bes[i + d * j] = s;
    }
  if (bes_out) memcpy(bes_out, bes, sizeof(double) * d * d);
  mo_logm(d, bes, lnb);
  /* tau* = 0.5 * Hooke(ln be*) */
  mo_hooke_stress(d, lam, mu, lnb, ts);
  for (int k = 0; k < d * d; ++k) ts[k] *= 0.5;
  double tr = 0;
  for (int i = 0; i < d; ++i) tr += ts[i + d * i];
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) td[i + d * j] = ts[i + d * j] - (i == j ? tr / d : 0.0);
  double ss = 0;
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) ss += td[i + d * j] * td[j + d * i];
  const double taueq = sqrt(3 * .5 * ss);
  for (int k = 0; k < d * d; ++k) Ns[k] = 3 * .5 * td[k] / taueq;
  double phi = taueq - tau_y0 - H * eps_old;
  phi = (phi < 0.0) ? 0.0 : phi; /* std::max(phi, 0.) keeps a NaN */
  const double dg = phi / (H + 3 * mu);
  for (int k = 0; k < d * d; ++k) tau[k] = ts[k] - 2 * dg * mu * Ns[k];
  /* updates */
  memcpy(F_prev_cur, F, sizeof(double) * d * d);
  for (int k = 0; k < d * d; ++k) lnb[k] -= 2 * dg * Ns[k];
  mo_expm(d, lnb, be_cur);
  *eps_cur += dg;
  /* tangent */
  if (phi > 0) {
    const double a0 = dg * mu / taueq;
    const double a1 = mu / (H + 3 * mu);
    for (int l = 0; l < d; ++l)
      for (int k = 0; k < d; ++k)
        for (int j = 0; j < d; ++j)
          for (int i = 0; i < d; ++i) {
            const double itrac = (i == j && k == l) ? 1.0 : 0.0;
            const double isym = 0.5 * (((i == k && j == l) ? 1.0 : 0.0) +
                                       ((i == l && j == k) ? 1.0 : 0.0));
            T4(C, i, j, k, l, d) = (Kb / 2. - mu / 3 + a0 * mu) * itrac +
                                   (1 - 3 * a0) * mu * isym +
                                   2 * mu * (a0 - a1) * Ns[i + d * j] * Ns[k + d * l];
          }
  } else {
    mo_hooke_C(d, lam, mu, C);
    for (int k = 0; k < d * d * d * d; ++k) C[k] *= 0.5;
  }
}

void mo_jacobi_sym3(const double *A, double *ev, double *V);

/*
 * Consistent tangent K(i,m,j,n) = dP_im/dF_jn of the HEP1 stress update
 * (no reference counterpart: the reference's coded tangent,
 * materials_toolbox.hh:394-406, is not the derivative of P and makes CG
 * diverge).  With be* = F A F^T (A fixed by the history), W = F^-1 be*,
 * dbe* = dF W + W^T dF^T, d(ln be*) = L : dbe* (Daleckii-Krein,
 * L_ijkl = sum_mn g_mn v_mi v_nj v_mk v_nl, g_mn = (ln l_n - ln l_m)/(l_n - l_m),
 * g_mm = 1/l_m), dtau = C : d(ln be*) with C the algorithmic modulus of
 * the return mapping (the reference's C, material_hyper_elasto_plastic1.hh:
 * 284-309), and P = tau F^-T:
 *   K(i,m,j,n) = sum_k [sum_q (DL_ikjq + DL_ikqj) W_nq] Finv_mk - P_in Finv_mj,
 *   DL = C : L.
 * C is the dim^2 x dim^2 T4 matrix returned by mo_hep1_eval_kirchhoff.
 */
void mo_hep1_consistent_tangent(int d, const double *F, const double *bes,
                                const double *C, const double *P, double *K) {
  const int nc = d * d;
  double Fi[9], W[9], ev[3], V[9], L[81], DL[81];
  mo_inverse(d, F, Fi);
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      double s = 0;
      for (int k = 0; k < d; ++k) s += Fi[i + d * k] * bes[k + d * j];
      W[i + d * j] = s;
    }
  if (d == 3) {
    mo_jacobi_sym3(bes, ev, V);
  } else {
    /* 2x2: rotation angle in closed form */
    const double a = bes[0], b = 0.5 * (bes[2] + bes[1]), c = bes[3];
    const double th = 0.5 * atan2(2 * b, a - c);
    const double cs = cos(th), sn = sin(th);
    V[0] = cs; V[1] = sn; V[2] = -sn; V[3] = cs;
    ev[0] = cs * cs * a + 2 * cs * sn * b + sn * sn * c;
    ev[1] = sn * sn * a - 2 * cs * sn * b + cs * cs * c;
  }
  double g[9];
  for (int m = 0; m < d; ++m)
    for (int n = 0; n < d; ++n) {
      const double h = ev[n] - ev[m];
      g[m + d * n] = h == 0.0 ? 1.0 / ev[m] : log1p(h / ev[m]) / h;
    }
  for (int l = 0; l < d; ++l)
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) {
          double s = 0;
          for (int m = 0; m < d; ++m)
            for (int n = 0; n < d; ++n)
              s += g[m + d * n] * V[i + d * m] * V[j + d * n] * V[k + d * m] * V[l + d * n];
          L[(i + d * j) + nc * (k + d * l)] = s;
        }
  /* DL = C : L as nc x nc matrices */
  for (int c = 0; c < nc; ++c)
    for (int r = 0; r < nc; ++r) {
      double s = 0;
      for (int q = 0; q < nc; ++q) s += C[r + nc * q] * L[q + nc * c];
      DL[r + nc * c] = s;
    }
  for (int n = 0; n < d; ++n)
    for (int j = 0; j < d; ++j)
      for (int m = 0; m < d; ++m)
        for (int i = 0; i < d; ++i) {
          double s = 0;
          for (int k = 0; k < d; ++k) {
            double e = 0;
            for (int q = 0; q < d; ++q)
              e += (DL[(i + d * k) + nc * (j + d * q)] + DL[(i + d * k) + nc * (q + d * j)]) *
                   W[n + d * q];
            s += e * Fi[m + d * k];
          }
          K[(i + d * m) + nc * (j + d * n)] = s - P[i + d * n] * Fi[m + d * j];
        }
}
