/*
 * Projections, cell, SolverCG and the Newton drivers of the oracle
 * (test infrastructure only).  Control flow is a literal restatement of
 *   src/fft/projection_small_strain.cc:42-86, projection_default.cc:53-63,
 *   src/fft/projection_finite_strain_fast.cc:43-87, projection_finite_strain.cc:61-95,
 *   src/cell/cell_base.cc:129-208,445-526,
 *   src/materials/material_muSpectre_base.hh:272-401,
 *   src/solver/solver_cg.cc:54-110, src/solver/solvers.cc:44-408.
 */
#include "muspectre_oracle.h"
#include "mo_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MO_MAXMAT 64

void mo_inverse(int d, const double *A, double *inv);
double mo_lambda(double E, double nu);
double mo_mu(double E, double nu);
void mo_hooke_C(int d, double lam, double mu, double *C);
void mo_hooke_stress(int d, double lam, double mu, const double *E, double *S);
void mo_tensmult(int d, const double *C, const double *B, double *R);
void mo_green_lagrange(int d, const double *F, double *E);

typedef struct {
  int family;
  double young, poisson, lambda, mu, tau_y0, H;
  double C[81];
  long n, cap;
  long *pix;
  int ppn;    /* per-pixel doubles */
  double *pp; /* LE2: eigenstrain; LE3: C; LE4: (lambda, mu) */
  /* HEP1 state fields: two buffers each; buffer `cur` is current() */
  double *Fp[2], *be[2], *ep[2];
  int cur;
} mo_material;

struct mo_cell {
  int dim, form, nc, allow_even, initialised;
  int res[3], hres[3];
  double L[3];
  long N, kpts;
  int nmat;
  mo_material mats[MO_MAXMAT];
  double *F, *P, *K;
  double *tmp, *cg_r, *cg_p, *cg_Ap;
  mo_cpx *work;
  double *freqs[3];
};

/* ---- projection operator at one k-point --------------------------------- */
static void kcoord(int d, const int *h, long k, int *c) {
  for (int a = d - 1; a >= 0; --a) {
    c[a] = (int)(k % h[a]);
    k /= h[a];
  }
}

/* unit wave vector (src/fft/fft_utils.hh:109-129); returns 0 where the
 * operator is zero: k = 0 (projection_*.cc: Ghat[0].setZero) or, with the
 * even-grid policy, any Nyquist index on an even axis */
static int unit_xi(int d, const int *res, double *const *freqs, const int *c,
                   int allow_even, long k, double *xi) {
  double n2 = 0;
  for (int a = 0; a < d; ++a) {
    xi[a] = freqs[a][c[a]];
    n2 += xi[a] * xi[a];
  }
  if (k == 0) return 0;
  if (allow_even)
    for (int a = 0; a < d; ++a)
      if (res[a] % 2 == 0 && c[a] == res[a] / 2) return 0;
  const double nrm = sqrt(n2);
  for (int a = 0; a < d; ++a) xi[a] /= nrm;
  return 1;
}

/* projection_small_strain.cc:61-80: the 81 (or 16) entries of Ghat */
static void ghat_small(int d, const double *xi, double *G) {
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j)
      for (int l = 0; l < d; ++l)
        for (int m = 0; m < d; ++m) {
          const double kjl = j == l, kjm = j == m, kil = i == l, kim = i == m;
          T4(G, i, j, l, m, d) = 0.5 * (xi[i] * kjl * xi[m] + xi[i] * kjm * xi[l] +
                                        xi[j] * kil * xi[m] + xi[j] * kim * xi[l]) -
                                 xi[i] * xi[j] * xi[l] * xi[m];
        }
}

/* projection_finite_strain.cc:73: outer_under(I, xi xi^T): G(i,j,k,l) = d_ik xi_j xi_l */
static void ghat_finite(int d, const double *xi, double *G) {
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j)
      for (int k = 0; k < d; ++k)
        for (int l = 0; l < d; ++l) T4(G, i, j, k, l, d) = (i == k) * xi[j] * xi[l];
}

static int check_sizes(int d, const int *res, int allow_even) {
  if (allow_even) return MO_OK;
  for (int a = 0; a < d; ++a)
    if (res[a] % 2 == 0) {
      mo_set_error("Only an odd number of gridpoints in each direction is supported");
      return MO_PROJECTION_ERROR;
    }
  return MO_OK;
}

static double **make_freqs(int d, const int *res, const double *L) {
  double **f = (double **)calloc(3, sizeof(double *));
  for (int a = 0; a < d; ++a) {
    f[a] = (double *)malloc(sizeof(double) * res[a]);
    mo_fft_freqs(res[a], L[a], f[a]);
  }
  return f;
}
static void free_freqs(double **f) {
  for (int a = 0; a < 3; ++a) free(f[a]);
  free(f);
}

int mo_projection_operator(int kind, int d, const int *res, const double *L,
                           int allow_even, double *op) {
  int e = check_sizes(d, res, allow_even);
  if (e) return e;
  int h[3];
  mo_hermitian_sizes(d, res, h);
  const long kpts = mo_grid_size(d, h);
  double **fr = make_freqs(d, res, L);
  const int nc = d * d, nop = kind == MO_PROJ_FINITE_STRAIN_FAST ? d : nc * nc;
  for (long k = 0; k < kpts; ++k) {
    int c[3];
    double xi[3];
    kcoord(d, h, k, c);
    double *o = op + k * nop;
    if (!unit_xi(d, res, fr, c, allow_even, k, xi)) {
      for (int q = 0; q < nop; ++q) o[q] = 0.0;
      continue;
    }
    if (kind == MO_PROJ_SMALL_STRAIN) ghat_small(d, xi, o);
    else if (kind == MO_PROJ_FINITE_STRAIN) ghat_finite(d, xi, o);
    else
      for (int a = 0; a < d; ++a) o[a] = xi[a];
  }
  free_freqs(fr);
  return MO_OK;
}

/* per-k-point application; f is nc complex values */
static void project_k(int kind, int d, const double *xi, double norm, mo_cpx *f) {
  const int nc = d * d;
  mo_cpx g[9];
  if (kind == MO_PROJ_FINITE_STRAIN_FAST) {
    /* projection_finite_strain_fast.cc:81-85: f <- norm (f xi) xi^T */
    for (int i = 0; i < d; ++i) {
      mo_cpx s = {0, 0};
      for (int k = 0; k < d; ++k) {
        s.re += f[i + d * k].re * xi[k];
        s.im += f[i + d * k].im * xi[k];
      }
      for (int j = 0; j < d; ++j) {
        g[i + d * j].re = norm * (s.re * xi[j]);
        g[i + d * j].im = norm * (s.im * xi[j]);
      }
    }
  } else {
    /* projection_default.cc:57-61: f <- norm (G f) */
    double G[81];
    if (kind == MO_PROJ_SMALL_STRAIN) ghat_small(d, xi, G);
    else ghat_finite(d, xi, G);
    for (int r = 0; r < nc; ++r) {
      mo_cpx s = {0, 0};
      for (int c = 0; c < nc; ++c) {
        s.re += G[r + nc * c] * f[c].re;
        s.im += G[r + nc * c] * f[c].im;
      }
      g[r].re = norm * s.re;
      g[r].im = norm * s.im;
    }
  }
  for (int r = 0; r < nc; ++r) f[r] = g[r];
}

static int apply_projection_ws(int kind, int d, const int *res, double *const *fr,
                               int allow_even, double *field, mo_cpx *work) {
  int h[3];
  mo_hermitian_sizes(d, res, h);
  const long kpts = mo_grid_size(d, h);
  const int nc = d * d;
  const double norm = 1.0 / (double)mo_grid_size(d, res);
  int e = mo_fft_r2c(d, res, nc, field, (double *)work);
  if (e) return e;
#pragma omp parallel for schedule(static) num_threads(mo_get_threads())
  for (long k = 0; k < kpts; ++k) {
    int c[3];
    double xi[3];
    kcoord(d, h, k, c);
    mo_cpx *f = work + k * nc;
    if (!unit_xi(d, res, fr, c, allow_even, k, xi)) {
      for (int q = 0; q < nc; ++q) f[q].re = f[q].im = 0.0;
      continue;
    }
    project_k(kind, d, xi, norm, f);
  }
  return mo_fft_c2r(d, res, nc, (const double *)work, field);
}

int mo_apply_projection(int kind, int d, const int *res, const double *L,
                        int allow_even, double *field) {
  int e = check_sizes(d, res, allow_even);
  if (e) return e;
  int h[3];
  mo_hermitian_sizes(d, res, h);
  const long kpts = mo_grid_size(d, h);
  mo_cpx *w = (mo_cpx *)malloc(sizeof(mo_cpx) * (size_t)(kpts * d * d));
  if (!w) return MO_ALLOC_ERROR;
  double **fr = make_freqs(d, res, L);
  e = apply_projection_ws(kind, d, res, fr, allow_even, field, w);
  free_freqs(fr);
  free(w);
  return e;
}

/* ---- cell ---------------------------------------------------------------- */
int mo_cell_create(int d, const int *res, const double *L, int form,
                   int allow_even, mo_cell **out) {
  *out = NULL;
  if (d != 2 && d != 3) {
    mo_set_error("only 2D and 3D cells");
    return MO_RUNTIME_ERROR;
  }
  int e = check_sizes(d, res, allow_even);
  if (e) return e;
  mo_cell *c = (mo_cell *)calloc(1, sizeof(mo_cell));
  c->dim = d;
  c->form = form;
  c->nc = d * d;
  c->allow_even = allow_even;
  for (int a = 0; a < d; ++a) {
    c->res[a] = res[a];
    c->L[a] = L[a];
  }
  mo_hermitian_sizes(d, res, c->hres);
  c->N = mo_grid_size(d, res);
  c->kpts = mo_grid_size(d, c->hres);
  c->F = (double *)calloc((size_t)(c->N * c->nc), sizeof(double));
  c->P = (double *)calloc((size_t)(c->N * c->nc), sizeof(double));
  c->tmp = (double *)calloc((size_t)(c->N * c->nc), sizeof(double));
  c->work = (mo_cpx *)calloc((size_t)(c->kpts * c->nc), sizeof(mo_cpx));
  for (int a = 0; a < d; ++a) {
    c->freqs[a] = (double *)malloc(sizeof(double) * res[a]);
    mo_fft_freqs(res[a], L[a], c->freqs[a]);
  }
  if (!c->F || !c->P || !c->tmp || !c->work) {
    mo_cell_destroy(c);
    return MO_ALLOC_ERROR;
  }
  *out = c;
  return MO_OK;
}

void mo_cell_destroy(mo_cell *c) {
  if (!c) return;
  for (int m = 0; m < c->nmat; ++m) {
    mo_material *mt = &c->mats[m];
    free(mt->pix);
    free(mt->pp);
    for (int b = 0; b < 2; ++b) {
      free(mt->Fp[b]);
      free(mt->be[b]);
      free(mt->ep[b]);
    }
  }
  free(c->F); free(c->P); free(c->K); free(c->tmp);
  free(c->cg_r); free(c->cg_p); free(c->cg_Ap);
  free(c->work);
  for (int a = 0; a < 3; ++a) free(c->freqs[a]);
  free(c);
}

long mo_cell_nb_pixels(const mo_cell *c) { return c->N; }
int mo_cell_nb_components(const mo_cell *c) { return c->nc; }
double *mo_cell_strain(mo_cell *c) { return c->F; }
double *mo_cell_stress(mo_cell *c) { return c->P; }
double *mo_cell_tangent(mo_cell *c) { return c->K; }

int mo_cell_add_material(mo_cell *c, int family, double young, double poisson,
                         double tau_y0, double H) {
  if (c->nmat >= MO_MAXMAT) return -MO_RUNTIME_ERROR;
  if ((family == MO_HEP1 || family == MO_HEP1_CT) && c->form != MO_FINITE_STRAIN) {
    mo_set_error("HyperElastoPlastic1 needs the finite-strain formulation");
    return -MO_RUNTIME_ERROR;
  }
  mo_material *m = &c->mats[c->nmat];
  memset(m, 0, sizeof(*m));
  m->family = family;
  m->young = young;
  m->poisson = poisson;
  m->tau_y0 = tau_y0;
  m->H = H;
  const int d = c->dim;
  if (family == MO_LE1 || family == MO_LE2 || family == MO_HEP1 || family == MO_HEP1_CT) {
    m->lambda = mo_lambda(young, poisson);
    m->mu = mo_mu(young, poisson);
    mo_hooke_C(d, m->lambda, m->mu, m->C);
  }
  m->ppn = family == MO_LE2 ? c->nc : family == MO_LE3 ? c->nc * c->nc
          : family == MO_LE4 ? 2 : 0;
  return c->nmat++;
}

/* material_linear_elastic_generic.cc:46-65: C(i,j,k,l) = C_voigt(v(i,j), v(k,l))
 * with v = VoigtConversion<dim>::sym_mat (voigt_conversion.hh:118-125) */
int mo_cell_add_material_generic(mo_cell *c, const double *C_voigt) {
  static const int v2[2][2] = {{0, 2}, {2, 1}};
  static const int v3[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
  const int d = c->dim, nc = c->nc, vs = d == 2 ? 3 : 6;
  int mi = mo_cell_add_material(c, MO_LEG, 0.0, 0.0, 0.0, 0.0);
  if (mi < 0) return mi;
  mo_material *m = &c->mats[mi];
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j)
      for (int k = 0; k < d; ++k)
        for (int l = 0; l < d; ++l) {
          const int a = d == 2 ? v2[i][j] : v3[i][j], b = d == 2 ? v2[k][l] : v3[k][l];
          m->C[(i + d * j) + nc * (k + d * l)] = C_voigt[a * vs + b];
        }
  return mi;
}

/* material_base.cc:56-58 and the LE2-4 add_pixel overloads */
int mo_material_add_pixel(mo_cell *c, int mi, long pixel, const double *params) {
  if (mi < 0 || mi >= c->nmat || pixel < 0 || pixel >= c->N) {
    mo_set_error("bad material or pixel");
    return MO_RUNTIME_ERROR;
  }
  mo_material *m = &c->mats[mi];
  if (m->ppn && !params) {
    mo_set_error(m->family == MO_LE2
                     ? "this material needs pixels with and eigenstrain"
                     : "this material needs pixels with Youngs modulus and Poisson ratio.");
    return MO_RUNTIME_ERROR;
  }
  if (m->n == m->cap) {
    long nc = m->cap ? 2 * m->cap : 64;
    m->pix = (long *)realloc(m->pix, sizeof(long) * nc);
    if (m->ppn) m->pp = (double *)realloc(m->pp, sizeof(double) * nc * m->ppn);
    m->cap = nc;
  }
  m->pix[m->n] = pixel;
  if (m->family == MO_LE2) {
    memcpy(m->pp + m->n * m->ppn, params, sizeof(double) * m->ppn);
  } else if (m->family == MO_LE3) {
    /* material_linear_elastic3.cc:57-68: C = compute_C(lambda, mu) */
    mo_hooke_C(c->dim, mo_lambda(params[0], params[1]), mo_mu(params[0], params[1]),
               m->pp + m->n * m->ppn);
  } else if (m->family == MO_LE4) {
    /* material_linear_elastic4.cc:62-73 */
    m->pp[m->n * 2 + 0] = mo_lambda(params[0], params[1]);
    m->pp[m->n * 2 + 1] = mo_mu(params[0], params[1]);
  }
  m->n++;
  return MO_OK;
}

/* cell_base.cc:445-454 with the coverage check of :486-526 and the HEP1
 * initialise of material_hyper_elasto_plastic1.cc:76-84 */
int mo_cell_initialise(mo_cell *c) {
  if (c->initialised) return MO_OK;
  int *assign = (int *)malloc(sizeof(int) * (size_t)c->N);
  for (long i = 0; i < c->N; ++i) assign[i] = -1;
  for (int mi = 0; mi < c->nmat; ++mi) {
    mo_material *m = &c->mats[mi];
    for (long q = 0; q < m->n; ++q) {
      long p = m->pix[q];
      if (assign[p] >= 0) {
        mo_set_error("Pixel %ld is already assigned to material %d and cannot be reassigned to material %d", p, assign[p], mi);
        free(assign);
        return MO_RUNTIME_ERROR;
      }
      assign[p] = mi;
    }
  }
  for (long i = 0; i < c->N; ++i)
    if (assign[i] < 0) {
      mo_set_error("The following pixels have were not assigned a material: %ld, ...", i);
      free(assign);
      return MO_RUNTIME_ERROR;
    }
  free(assign);
  const int nc = c->nc, d = c->dim;
  for (int mi = 0; mi < c->nmat; ++mi) {
    mo_material *m = &c->mats[mi];
    if (m->family != MO_HEP1 && m->family != MO_HEP1_CT) continue;
    for (int b = 0; b < 2; ++b) {
      m->Fp[b] = (double *)calloc((size_t)(m->n * nc), sizeof(double));
      m->be[b] = (double *)calloc((size_t)(m->n * nc), sizeof(double));
      m->ep[b] = (double *)calloc((size_t)m->n, sizeof(double));
    }
    m->cur = 0;
    for (long q = 0; q < m->n; ++q)
      for (int i = 0; i < d; ++i) {
        m->Fp[0][q * nc + i * (d + 1)] = 1.0;
        m->be[0][q * nc + i * (d + 1)] = 1.0;
      }
    m->cur = 1; /* save_history_variables(): cycle */
  }
  c->initialised = 1;
  return MO_OK;
}

int mo_cell_save_history(mo_cell *c) {
  for (int mi = 0; mi < c->nmat; ++mi)
    if (c->mats[mi].family == MO_HEP1 || c->mats[mi].family == MO_HEP1_CT) c->mats[mi].cur ^= 1;
  return MO_OK;
}

void mo_pk1_from_kirchhoff(int d, const double *F, const double *tau,
                           const double *C, double *P, double *K);
void mo_pk1_from_pk2(int d, const double *F, const double *S, const double *C,
                     double *P, double *K);

/* material_muSpectre_base.hh:292-401: per material, per own pixel */
static int eval_materials(mo_cell *c, int need_tangent) {
  const int d = c->dim, nc = c->nc, nt = nc * nc;
  if (!c->K) {
    c->K = (double *)calloc((size_t)(c->N * nt), sizeof(double));
    if (!c->K) return MO_ALLOC_ERROR;
  }
  for (int mi = 0; mi < c->nmat; ++mi) {
    mo_material *m = &c->mats[mi];
    const int cur = m->cur, old = cur ^ 1;
#pragma omp parallel for schedule(static) num_threads(mo_get_threads())
    for (long q = 0; q < m->n; ++q) {
      const long p = m->pix[q];
      const double *F = c->F + p * nc;
      double *P = c->P + p * nc;
      double Kl[81];
      double *K = need_tangent ? c->K + p * nt : Kl;
      double strain[9], S[9], Cl[81];
      const double *C = m->C;
      double lam = m->lambda, mu = m->mu;
      switch (m->family) {
        case MO_LE1:
        case MO_LE2: {
          /* LE1: Hooke on the strain (small) or Green-Lagrange (finite);
           * LE2: same on (strain - eigenstrain), material_linear_elastic2.hh:178-201 */
          if (c->form == MO_SMALL_STRAIN) memcpy(strain, F, sizeof(double) * nc);
          else mo_green_lagrange(d, F, strain);
          if (m->family == MO_LE2)
            for (int k = 0; k < nc; ++k) strain[k] -= m->pp[q * nc + k];
          mo_hooke_stress(d, lam, mu, strain, S);
          break;
        }
        case MO_LEG: {
          /* material_linear_elastic_generic.hh:164-176: S = C : E */
          if (c->form == MO_SMALL_STRAIN) memcpy(strain, F, sizeof(double) * nc);
          else mo_green_lagrange(d, F, strain);
          mo_tensmult(d, m->C, strain, S);
          break;
        }
        case MO_LE3:
        case MO_LE4: {
          /* material_linear_elastic3.hh:181-198 / _4.hh:189-210: C:eps */
          if (m->family == MO_LE3) {
            C = m->pp + q * nt;
          } else {
            mo_hooke_C(d, m->pp[2 * q], m->pp[2 * q + 1], Cl);
            C = Cl;
          }
          if (c->form == MO_SMALL_STRAIN) memcpy(strain, F, sizeof(double) * nc);
          else mo_green_lagrange(d, F, strain);
          mo_tensmult(d, C, strain, S);
          break;
        }
        case MO_HEP1: {
          double tau[9];
          mo_hep1_eval_kirchhoff(d, m->young, m->poisson, m->tau_y0, m->H, F,
                                 m->Fp[old] + q * nc, m->be[old] + q * nc,
                                 m->ep[old][q], m->Fp[cur] + q * nc,
                                 m->be[cur] + q * nc, &m->ep[cur][q], tau, Cl);
          mo_pk1_from_kirchhoff(d, F, tau, Cl, P, K);
          continue;
        }
        case MO_HEP1_CT: {
          double tau[9], bes[9];
          mo_hep1_eval_kirchhoff_bes(d, m->young, m->poisson, m->tau_y0, m->H, F,
                                     m->Fp[old] + q * nc, m->be[old] + q * nc,
                                     m->ep[old][q], m->Fp[cur] + q * nc,
                                     m->be[cur] + q * nc, &m->ep[cur][q], tau, Cl, bes);
          mo_pk1_from_kirchhoff(d, F, tau, Cl, P, K);
          if (need_tangent) mo_hep1_consistent_tangent(d, F, bes, Cl, P, K);
          continue;
        }
      }
      if (c->form == MO_SMALL_STRAIN) {
        memcpy(P, S, sizeof(double) * nc);
        memcpy(K, C, sizeof(double) * nt);
      } else {
        mo_pk1_from_pk2(d, F, S, C, P, K);
      }
    }
  }
  return MO_OK;
}

/* cell_base.cc:129-146 */
int mo_cell_evaluate_stress_tangent(mo_cell *c) {
  int e = mo_cell_initialise(c);
  if (e) return e;
  return eval_materials(c, 1);
}

/* cell_base.cc:112-124 (stress only; tangent scratch is discarded) */
int mo_cell_evaluate_stress(mo_cell *c) {
  int e = mo_cell_initialise(c);
  if (e) return e;
  return eval_materials(c, 0);
}

static int proj_kind(const mo_cell *c) {
  return c->form == MO_SMALL_STRAIN ? MO_PROJ_SMALL_STRAIN : MO_PROJ_FINITE_STRAIN_FAST;
}

/* cell_base.cc:201-208 */
int mo_cell_apply_projection(mo_cell *c, double *v) {
  return apply_projection_ws(proj_kind(c), c->dim, c->res, c->freqs,
                             c->allow_even, v, c->work);
}

/* cell_base.cc:182-188: dP = tensmult(K, dF) per voxel */
int mo_cell_tangent_product(mo_cell *c, const double *dF, double *dP) {
  if (!c->K) {
    mo_set_error("currently only implemented for cases where a stiffness matrix exists");
    return MO_RUNTIME_ERROR;
  }
  const int d = c->dim, nc = c->nc, nt = nc * nc;
#pragma omp parallel for schedule(static) num_threads(mo_get_threads())
  for (long p = 0; p < c->N; ++p) mo_tensmult(d, c->K + p * nt, dF + p * nc, dP + p * nc);
  return MO_OK;
}

/* cell_base.cc:150-192: K:dF, then project */
int mo_cell_directional_stiffness(mo_cell *c, const double *dF, double *dP) {
  int e = mo_cell_tangent_product(c, dF, dP);
  if (e) return e;
  return mo_cell_apply_projection(c, dP);
}

int mo_cell_get_internal(mo_cell *c, const char *name, int which, double *out,
                         int *nb_comp) {
  int found = 0, ncomp = 0;
  for (int mi = 0; mi < c->nmat; ++mi) {
    mo_material *m = &c->mats[mi];
    const int b = which == 0 ? m->cur : (m->cur ^ 1);
    const double *src = NULL;
    int k = 0;
    if ((m->family == MO_HEP1 || m->family == MO_HEP1_CT) && m->Fp[0]) {
      if (!strcmp(name, "eps_p")) { src = m->ep[b]; k = 1; }
      else if (!strcmp(name, "F_prev")) { src = m->Fp[b]; k = c->nc; }
      else if (!strcmp(name, "be")) { src = m->be[b]; k = c->nc; }
    } else if (m->family == MO_LE2 && !strcmp(name, "eigenstrain")) {
      src = m->pp; k = c->nc;
    }
    if (!src) continue;
    if (!found) {
      ncomp = k;
      for (long i = 0; i < c->N * k; ++i) out[i] = 0.0;
    }
    found = 1;
    for (long q = 0; q < m->n; ++q)
      memcpy(out + m->pix[q] * k, src + q * k, sizeof(double) * k);
  }
  if (!found) {
    mo_set_error("internal field %s does not exist in any material", name);
    return MO_RUNTIME_ERROR;
  }
  *nb_comp = ncomp;
  return MO_OK;
}

/* ---- SolverCG ------------------------------------------------------------ */
static double dot(const double *a, const double *b, long n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static) num_threads(mo_get_threads())
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static int ensure_cg(mo_cell *c) {
  const long n = c->N * c->nc;
  if (!c->cg_r) {
    c->cg_r = (double *)calloc((size_t)n, sizeof(double));
    c->cg_p = (double *)calloc((size_t)n, sizeof(double));
    c->cg_Ap = (double *)calloc((size_t)n, sizeof(double));
  }
  return (c->cg_r && c->cg_p && c->cg_Ap) ? MO_OK : MO_ALLOC_ERROR;
}

/* stopping-test records of mo_cg_solve (diagnostics, see the header) */
#define MO_CG_RECORDS_MAX 4096
static double g_cg_rec[MO_CG_RECORDS_MAX][4];
static int g_cg_nrec = 0;

int mo_cg_records(double *out, int cap) {
  const int n = g_cg_nrec < cap ? g_cg_nrec : cap;
  memcpy(out, g_cg_rec, sizeof(double) * 4 * (size_t)(n > 0 ? n : 0));
  return g_cg_nrec;
}

void mo_cg_records_clear(void) { g_cg_nrec = 0; }

static void cg_record(int iters, double rdr, double rdr_prev, double tol2) {
  if (g_cg_nrec >= MO_CG_RECORDS_MAX) return;
  g_cg_rec[g_cg_nrec][0] = iters;
  g_cg_rec[g_cg_nrec][1] = rdr;
  g_cg_rec[g_cg_nrec][2] = rdr_prev;
  g_cg_rec[g_cg_nrec][3] = tol2;
  ++g_cg_nrec;
}

/* src/solver/solver_cg.cc:54-110 */
int mo_cg_solve(mo_cell *c, double tol, int maxiter, const double *rhs,
                double *x, long *counter) {
  int e = ensure_cg(c);
  if (e) return e;
  const long n = c->N * c->nc;
  double *r = c->cg_r, *p = c->cg_p, *Ap = c->cg_Ap;
  for (long i = 0; i < n; ++i) x[i] = 0.0;
  /* r = A x - b: one full operator application (on x = 0) */
  e = mo_cell_directional_stiffness(c, x, r);
  if (e) return e;
  for (long i = 0; i < n; ++i) r[i] -= rhs[i];
  for (long i = 0; i < n; ++i) p[i] = -r[i];
  double rdr = dot(r, r, n);
  const double rhs_norm2 = dot(rhs, rhs, n);
  const double tol2 = tol * tol * rhs_norm2;
  double rdr_prev = rdr;
  int its = 0;
  for (int i = 0; i < maxiter && (rdr > tol2 || i == 0); ++i, ++(*counter)) {
    e = mo_cell_directional_stiffness(c, p, Ap);
    if (e) return e;
    const double alpha = rdr / dot(p, Ap, n);
#pragma omp parallel for schedule(static) num_threads(mo_get_threads())
    for (long q = 0; q < n; ++q) {
      x[q] += alpha * p[q];
      r[q] += alpha * Ap[q];
    }
    const double new_rdr = dot(r, r, n);
    const double beta = new_rdr / rdr;
    rdr_prev = rdr;
    rdr = new_rdr;
    ++its;
#pragma omp parallel for schedule(static) num_threads(mo_get_threads())
    for (long q = 0; q < n; ++q) p[q] = -r[q] + beta * p[q];
  }
  cg_record(its, rdr, rdr_prev, tol2);
  if (rdr < tol2) return MO_OK;
  mo_set_error("Conjugate gradient has not converged. After %ld steps, the solver  FAILED with  |r|/|b| = %g, cg_tol = %g",
               *counter, sqrt(rdr / rhs_norm2), tol);
  return MO_CONVERGENCE_ERROR;
}

int mo_cg_fixed_iterations(mo_cell *c, const double *rhs, double *x, int iters) {
  int e = ensure_cg(c);
  if (e) return e;
  const long n = c->N * c->nc;
  double *r = c->cg_r, *p = c->cg_p, *Ap = c->cg_Ap;
  for (long i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = -rhs[i];
    p[i] = rhs[i];
  }
  double rdr = dot(r, r, n);
  for (int i = 0; i < iters; ++i) {
    e = mo_cell_directional_stiffness(c, p, Ap);
    if (e) return e;
    const double alpha = rdr / dot(p, Ap, n);
#pragma omp parallel for schedule(static) num_threads(mo_get_threads())
    for (long q = 0; q < n; ++q) {
      x[q] += alpha * p[q];
      r[q] += alpha * Ap[q];
    }
    const double new_rdr = dot(r, r, n);
    const double beta = new_rdr / rdr;
    rdr = new_rdr;
#pragma omp parallel for schedule(static) num_threads(mo_get_threads())
    for (long q = 0; q < n; ++q) p[q] = -r[q] + beta * p[q];
  }
  return MO_OK;
}

/* ---- Newton drivers ------------------------------------------------------ */
/* src/solver/solver_common.cc:39-43 with rel_tol = 1e-8 */
static int check_symmetry(int d, const double *M) {
  double an = 0, n = 0;
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      const double a = M[i + d * j] - M[j + d * i];
      an += a * a;
      n += M[i + d * j] * M[i + d * j];
    }
  an = sqrt(an);
  n = sqrt(n);
  return (1e-8 >= an / n) || (1e-8 >= n);
}

typedef struct {
  double incr_norm, grad_norm, stress_norm, newton_tol, equil_tol;
  int message, has_converged;
} conv_state;

/* the convergence_test lambda of solvers.cc:152-163 */
static int convergence_test(conv_state *s) {
  const int incr_test = s->incr_norm / s->grad_norm <= s->newton_tol;
  const int stress_test = s->stress_norm < s->equil_tol;
  if (incr_test) s->message = 1;
  else if (stress_test) s->message = 2;
  s->has_converged = incr_test || stress_test;
  return s->has_converged;
}

static int newton_init(mo_cell *c, int nsteps, const double *load_steps) {
  const int d = c->dim, nc = c->nc;
  int e = mo_cell_initialise(c);
  if (e) return e;
  for (long p = 0; p < c->N; ++p)
    for (int k = 0; k < nc; ++k)
      c->F[p * nc + k] = (c->form == MO_FINITE_STRAIN && k % (d + 1) == 0) ? 1.0 : 0.0;
  if (c->form == MO_SMALL_STRAIN)
    for (int s = 0; s < nsteps; ++s)
      if (!check_symmetry(d, load_steps + s * nc)) {
        mo_set_error("all Δε must be symmetric!");
        return MO_SOLVER_ERROR;
      }
  return MO_OK;
}

static double norm2(const double *v, long n) { return sqrt(dot(v, v, n)); }

static int driver(mo_cell *c, int de_geus, int nsteps, const double *load_steps,
                  double cg_tol, int maxiter, double newton_tol,
                  double equil_tol, long *counter, mo_result *results,
                  double *grads, double *stresses) {
  const int nc = c->nc;
  const long n = c->N * nc;
  int e = newton_init(c, nsteps, load_steps);
  if (e) return e;
  double *incrF = (double *)calloc((size_t)n, sizeof(double));
  double *rhs = (double *)calloc((size_t)n, sizeof(double));
  double *DeltaF = de_geus ? (double *)calloc((size_t)n, sizeof(double)) : NULL;
  double prev[9] = {0};
  for (int s = 0; s < nsteps && !e; ++s) {
    const double *macro = load_steps + s * nc;
    if (!de_geus)
      for (long p = 0; p < c->N; ++p)
        for (int k = 0; k < nc; ++k) c->F[p * nc + k] += macro[k] - prev[k];
    conv_state cs = {2 * newton_tol, 1.0, 2 * equil_tol, newton_tol, equil_tol, 0, 0};
    int newt_iter = 0;
    for (; newt_iter < maxiter && !cs.has_converged; ++newt_iter) {
      e = mo_cell_evaluate_stress_tangent(c);
      if (e) break;
      if (de_geus && newt_iter == 0) {
        /* solvers.cc:353-363 */
        for (long p = 0; p < c->N; ++p)
          for (int k = 0; k < nc; ++k) DeltaF[p * nc + k] = macro[k] - prev[k];
        e = mo_cell_directional_stiffness(c, DeltaF, rhs);
        if (e) break;
        for (long q = 0; q < n; ++q) rhs[q] = -rhs[q];
        cs.stress_norm = norm2(rhs, n);
        if (convergence_test(&cs)) break;
        e = mo_cg_solve(c, cg_tol, maxiter, rhs, incrF, counter);
        if (e) break;
        for (long q = 0; q < n; ++q) c->F[q] += DeltaF[q];
      } else {
        for (long q = 0; q < n; ++q) rhs[q] = -c->P[q];
        e = mo_cell_apply_projection(c, rhs);
        if (e) break;
        cs.stress_norm = norm2(rhs, n);
        if (convergence_test(&cs)) break;
        e = mo_cg_solve(c, cg_tol, maxiter, rhs, incrF, counter);
        if (e) break;
      }
      for (long q = 0; q < n; ++q) c->F[q] += incrF[q];
      cs.incr_norm = norm2(incrF, n);
      cs.grad_norm = norm2(c->F, n);
      convergence_test(&cs);
    }
    if (e) break;
    for (int k = 0; k < nc; ++k) prev[k] = macro[k];
    const int ok = convergence_test(&cs);
    if (results) {
      results[s].success = ok;
      results[s].status = ok;
      results[s].message = cs.message;
      results[s].nb_it = newt_iter;
      results[s].nb_fev = *counter;
    }
    if (grads) memcpy(grads + s * n, c->F, sizeof(double) * n);
    if (stresses) memcpy(stresses + s * n, c->P, sizeof(double) * n);
    mo_cell_save_history(c);
  }
  free(incrF);
  free(rhs);
  free(DeltaF);
  return e;
}

/* src/solver/solvers.cc:44-217 */
int mo_newton_cg(mo_cell *c, int nsteps, const double *load_steps, double cg_tol,
                 int maxiter, double newton_tol, double equil_tol, long *counter,
                 mo_result *results, double *grads, double *stresses) {
  return driver(c, 0, nsteps, load_steps, cg_tol, maxiter, newton_tol, equil_tol,
                counter, results, grads, stresses);
}

/* src/solver/solvers.cc:220-408 */
int mo_de_geus(mo_cell *c, int nsteps, const double *load_steps, double cg_tol,
               int maxiter, double newton_tol, double equil_tol, long *counter,
               mo_result *results, double *grads, double *stresses) {
  return driver(c, 1, nsteps, load_steps, cg_tol, maxiter, newton_tol, equil_tol,
                counter, results, grads, stresses);
}
