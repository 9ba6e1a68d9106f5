/*
 * muspectre_oracle.h — CPU restatement (FP64, plain C99) of the µSpectre
 * Newton-CG inner loop.
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke()
 * check of __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it, and only as the checker or the timed CPU baseline.  The GPU library
 * (muspectre_amd) never links, loads or falls back to it.
 *
 * Every function follows a reference file:line (paths relative to the
 * reference repository root).  Conventions (all from the reference):
 *   - pixel index: row-major, last axis fastest (src/common/ccoord_operations.hh:179-192)
 *   - fields are AoS: nc components per pixel, component of a 2nd-order
 *     tensor A(i,j) at i + dim*j (Eigen column major, src/common/field_typed.hh:119)
 *   - 4th-order tensors: T(i,j,k,l) at (i+dim*j) + dim^2*(k+dim*l)
 *     (src/common/T4_map_proxy.hh:83-105)
 *   - Fourier grid: (N0, .., N_{d-1}/2+1), row-major, AoS complex with nc
 *     components per k-point (src/fft/fftw_engine.cc:40-77)
 *   - both transforms unnormalised; 1/N folded into the projection
 *     (src/fft/fft_engine_base.cc:49, src/fft/projection_default.cc:56-60)
 *
 * Each function is checked by tests/test_oracle_*.py against the reference's
 * golden vectors (tests/golden/reference_vectors.json), closed forms,
 * invariants, finite differences or brute force.  The coded Kirchhoff->PK1
 * tangent (mo_pk1_from_kirchhoff), for which no reference vector exists, is
 * fixed uniquely by two identities on random inputs: K(i,m,j,n) does not
 * depend on j, and F K over its first index equals the sums of C over its
 * last index.  HEP1 over many steps is held, at every step of a
 * loading/unloading path, to what its return mapping fixes for any state
 * (deviatoric flow, Hencky stress of the new be, the yield surface, and
 * be = F F^T on an elastic path).  Parity unpinned (nothing of that kind
 * fixes them): logm/expm at exactly repeated eigenvalues against Eigen's
 * computeDirect rounding (the mathematical function is matched to 1e-13);
 * the reference's own Newton/CG iteration counts at the benchmark
 * configurations.
 */
#ifndef MUSPECTRE_ORACLE_H
#define MUSPECTRE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* error codes, mapped by the callers onto the reference exception types */
enum {
  MO_OK = 0,
  MO_CONVERGENCE_ERROR = 1, /* ConvergenceError, src/solver/solver_common.hh:84-91 */
  MO_SOLVER_ERROR = 2,      /* SolverError, src/solver/solver_common.hh:84-91     */
  MO_PROJECTION_ERROR = 3,  /* ProjectionError, src/fft/projection_base.hh:50-58  */
  MO_RUNTIME_ERROR = 4,     /* std::runtime_error (coverage, sizes, ...)          */
  MO_ALLOC_ERROR = 5
};

enum { MO_FINITE_STRAIN = 0, MO_SMALL_STRAIN = 1 }; /* src/common/common.hh:149-153 */

enum {
  MO_PROJ_SMALL_STRAIN = 0,       /* projection_small_strain.cc        */
  MO_PROJ_FINITE_STRAIN_FAST = 1, /* projection_finite_strain_fast.cc  */
  MO_PROJ_FINITE_STRAIN = 2       /* projection_finite_strain.cc (full Ghat) */
};

/* MO_HEP1_CT: the HEP1 stress update with the consistent tangent dP/dF
 * (no reference counterpart; see mo_hep1_consistent_tangent) */
/* MO_LEG: MaterialLinearElasticGeneric (material_linear_elastic_generic.hh),
 * one stiffness per material from a Voigt matrix */
enum {
  MO_LE1 = 1, MO_LE2 = 2, MO_LE3 = 3, MO_LE4 = 4, MO_HEP1 = 5, MO_HEP1_CT = 6, MO_LEG = 7
};

/* threads used by the OpenMP loops (1 = the reference's single-core path) */
void mo_set_threads(int n);
int mo_get_threads(void);
const char *mo_last_error(void);

/* ---- grid helpers --------------------------------------------------- */
long mo_grid_size(int dim, const int *res);
void mo_hermitian_sizes(int dim, const int *res, int *out);
/* numpy fftfreq(n)*n order, divided by length (src/fft/fft_utils.cc:41-56) */
void mo_fft_freqs(int n, double length, double *out);

/* ---- FFT (replaces FFTW r2c/c2r, src/fft/fftw_engine.cc:98-110,130-158) */
/* in: N*nc reals (AoS); out: Kpts*nc complex (interleaved re,im) */
int mo_fft_r2c(int dim, const int *res, int nc, const double *in, double *out);
/* in: Kpts*nc complex (not modified); out: N*nc reals, unnormalised */
int mo_fft_c2r(int dim, const int *res, int nc, const double *in, double *out);

/* ---- projections ---------------------------------------------------- */
/* allow_even=0 reproduces the reference's odd-only check; 1 applies the
 * even-grid Nyquist policy (operator zeroed on every Nyquist index). */
int mo_projection_operator(int kind, int dim, const int *res,
                           const double *lengths, int allow_even, double *op);
int mo_apply_projection(int kind, int dim, const int *res,
                        const double *lengths, int allow_even, double *field);

/* ---- small-tensor algebra (src/common/eigen_tools.hh:209-341) ---------- */
void mo_eig_sym3(const double *A, double *evals);            /* ascending */
void mo_eigsys_sym3(const double *A, double *evals, double *evecs);
void mo_eig_sym2(const double *A, double *evals);
void mo_logm(int dim, const double *A, double *out);
void mo_expm(int dim, const double *A, double *out);
/* literal reference forms, used by the tests to show agreement */
void mo_logm_sylvester(int dim, const double *A, double *out);
void mo_expm_eig(const double *A, double *out);
void mo_inverse(int dim, const double *A, double *out);

/* ---- single-point constitutive laws --------------------------------- */
/* LE1 finite or small strain; F in, P and K (dim^4) out */
void mo_le1_eval(int dim, int form, double young, double poisson,
                 const double *F, double *P, double *K);
/* HEP1: material-level (Kirchhoff stress and its tangent, before the PK1
 * conversion); state buffers are updated in place like the reference's
 * current() values.  F_prev_old, be_old, eps_old are the old() values. */
void mo_hep1_eval_kirchhoff(int dim, double young, double poisson,
                            double tau_y0, double H, const double *F,
                            const double *F_prev_old, const double *be_old,
                            double eps_old, double *F_prev_cur,
                            double *be_cur, double *eps_cur, double *tau,
                            double *C);
/* same, also returning the trial elastic left Cauchy-Green tensor be* */
void mo_hep1_eval_kirchhoff_bes(int dim, double young, double poisson,
                                double tau_y0, double H, const double *F,
                                const double *F_prev_old, const double *be_old,
                                double eps_old, double *F_prev_cur,
                                double *be_cur, double *eps_cur, double *tau,
                                double *C, double *bes);
/* consistent tangent dP/dF of the HEP1 update (no reference counterpart);
 * C is the algorithmic modulus from mo_hep1_eval_kirchhoff, P = tau F^-T */
void mo_hep1_consistent_tangent(int dim, const double *F, const double *bes,
                                const double *C, const double *P, double *K);
void mo_jacobi_sym3(const double *A, double *evals, double *evecs);
/* PK1 conversions (src/materials/materials_toolbox.hh:313-410) */
void mo_pk1_from_pk2(int dim, const double *F, const double *S,
                     const double *C, double *P, double *K);
void mo_pk1_from_kirchhoff(int dim, const double *F, const double *tau,
                           const double *C, double *P, double *K);

/* ---- cell / solver -------------------------------------------------- */
typedef struct mo_cell mo_cell;

int mo_cell_create(int dim, const int *res, const double *lengths, int form,
                   int allow_even, mo_cell **out);
void mo_cell_destroy(mo_cell *c);
long mo_cell_nb_pixels(const mo_cell *c);
int mo_cell_nb_components(const mo_cell *c);
/* returns material index >= 0, or <0 on error */
int mo_cell_add_material(mo_cell *c, int family, double young,
                         double poisson, double tau_y0, double H);
/* MaterialLinearElasticGeneric(name, C_voigt), material_linear_elastic_generic.cc:
 * 40-65: C_voigt is vsize x vsize row-major (3 in 2D: xx, yy, xy; 6 in 3D:
 * xx, yy, zz, yz, xz, xy, voigt_conversion.hh:118-125); returns the material
 * index >= 0, or <0 on error */
int mo_cell_add_material_generic(mo_cell *c, const double *C_voigt);
/* params: LE2 eigenstrain (dim*dim col-major); LE3/LE4 (Young, Poisson) */
int mo_material_add_pixel(mo_cell *c, int mat, long pixel,
                          const double *params);
int mo_cell_initialise(mo_cell *c);
double *mo_cell_strain(mo_cell *c);   /* F, N*nc */
double *mo_cell_stress(mo_cell *c);   /* P, N*nc */
double *mo_cell_tangent(mo_cell *c);  /* K, N*nc*nc (NULL before first eval) */
int mo_cell_evaluate_stress_tangent(mo_cell *c);
int mo_cell_evaluate_stress(mo_cell *c);
int mo_cell_directional_stiffness(mo_cell *c, const double *dF, double *dP);
/* K:dF per voxel only (no projection) */
int mo_cell_tangent_product(mo_cell *c, const double *dF, double *dP);
int mo_cell_apply_projection(mo_cell *c, double *v);
int mo_cell_save_history(mo_cell *c);
/* copy a material's internal state field (HEP1: "eps_p", "F_prev", "be";
 * LE2: "eigenstrain") into a global pixel field; which=0 current, 1 old */
int mo_cell_get_internal(mo_cell *c, const char *name, int which,
                         double *out, int *nb_comp);

/* SolverCG::solve (src/solver/solver_cg.cc:54-110); *counter is the
 * solver's cumulative counter and is updated. */
int mo_cg_solve(mo_cell *c, double tol, int maxiter, const double *rhs,
                double *x, long *counter);

/* Diagnostics of the stopping test (no reference counterpart): every
 * mo_cg_solve appends one record {iterations, r.r after the last iteration,
 * r.r after the one before (r.r at the start for a single iteration),
 * tol2}, so that a test can see how close r.r came to tol2 where the loop
 * stopped.  mo_cg_records copies up to cap records (4 doubles each) and
 * returns how many are held; mo_cg_records_clear empties the list (at most
 * 4096 records are kept). */
int mo_cg_records(double *out, int cap);
void mo_cg_records_clear(void);

typedef struct {
  int success;
  int status;
  int nb_it;
  long nb_fev;
  int message; /* 0 "Has not converged", 1 "Residual  tolerance reached",
                  2 "Reached stress divergence tolerance" */
} mo_result;

/* newton_cg / de_geus over nsteps load steps (dim x dim col-major each).
 * grads/stresses: nsteps*N*nc or NULL.  *counter: cumulative CG count. */
int mo_newton_cg(mo_cell *c, int nsteps, const double *load_steps,
                 double cg_tol, int maxiter, double newton_tol,
                 double equil_tol, long *counter, mo_result *results,
                 double *grads, double *stresses);
int mo_de_geus(mo_cell *c, int nsteps, const double *load_steps,
               double cg_tol, int maxiter, double newton_tol,
               double equil_tol, long *counter, mo_result *results,
               double *grads, double *stresses);

/* timing helper for the CPU baseline: run `iters` CG iterations of the
 * operator/update sequence on the current tangent (no convergence test). */
int mo_cg_fixed_iterations(mo_cell *c, const double *rhs, double *x, int iters);

#ifdef __cplusplus
}
#endif
#endif
