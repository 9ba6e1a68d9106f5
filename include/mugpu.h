/*
 * mugpu.h — C-ABI of the MI355X (gfx950) implementation of µSpectre's
 * Newton-CG inner loop.  Plain pointers, sizes and opaque handles only.
 *
 * Every entry point names the reference interface it stands in for
 * (paths relative to the reference repository root).  All functions return
 * 0 on success or one of the MUGPU_E* codes; mugpu_last_error() gives the
 * message.  The host layer maps the codes onto the reference's exception
 * types (ProjectionError, SolverError, ConvergenceError, runtime_error).
 *
 * Array conventions are the reference's: real fields are AoS with nc
 * components per pixel (tensor component A(i,j) at i + dim*j), pixels in
 * row-major order with the last axis fastest; spectra are AoS complex over
 * the Hermitian grid (N0, .., N_{d-1}/2+1)
 * (src/common/field_typed.hh:119, src/common/ccoord_operations.hh:179-192,
 *  src/fft/fftw_engine.cc:40-77).  Pointers named d_* are device pointers.
 *
 * precision: MUGPU_FP64 (plain path) or MUGPU_FP32 (operator stage in
 * single precision).  Threading: call the library from one thread at a
 * time.  All handles share the library's stream, its pinned result words,
 * its device scratch and its timing state, so calls that overlap from
 * several threads race, whatever the handles; the functions that return
 * host scalars synchronise the stream.  The text of mugpu_last_error() is
 * that of the calling thread's last failed call.
 */
#ifndef MUGPU_H
#define MUGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (src/solver/solver_common.hh:84-91,
 *      src/fft/projection_base.hh:50-58) ---------------------------------- */
#define MUGPU_OK 0
#define MUGPU_ECONVERGENCE 1 /* ConvergenceError */
#define MUGPU_ESOLVER 2      /* SolverError */
#define MUGPU_EPROJECTION 3  /* ProjectionError */
#define MUGPU_ERUNTIME 4     /* std::runtime_error */
#define MUGPU_EMEMORY 5      /* device allocation failed */
#define MUGPU_EHIP 6         /* HIP runtime error */
#define MUGPU_EARG 7         /* invalid argument (size/shape mismatch) */

#define MUGPU_FP64 0
#define MUGPU_FP32 1

/* src/common/common.hh:149-153 */
#define MUGPU_FINITE_STRAIN 0
#define MUGPU_SMALL_STRAIN 1

/* projections: projection_small_strain.cc, projection_finite_strain_fast.cc,
 * projection_finite_strain.cc */
#define MUGPU_PROJ_SMALL_STRAIN 0
#define MUGPU_PROJ_FINITE_STRAIN_FAST 1
#define MUGPU_PROJ_FINITE_STRAIN 2

/* material families: material_linear_elastic{1,2,3,4}.hh,
 * material_hyper_elasto_plastic1.hh */
#define MUGPU_LE1 1
#define MUGPU_LE2 2
#define MUGPU_LE3 3
#define MUGPU_LE4 4
#define MUGPU_HEP1 5
/* HEP1 stress update with the consistent tangent dP/dF (no reference
 * counterpart: the reference's coded HEP1 tangent, materials_toolbox.hh:
 * 394-406, is not the derivative of P and makes CG diverge) */
#define MUGPU_HEP1_CT 6
/* MaterialLinearElasticGeneric (material_linear_elastic_generic.hh): one
 * stiffness tensor per material, given in Voigt notation; added with
 * mugpu_cell_add_material_generic */
#define MUGPU_LEG 7

typedef struct mugpu_fft_s *mugpu_fft;
typedef struct mugpu_proj_s *mugpu_proj;
typedef struct mugpu_cell_s *mugpu_cell;

/* ---- runtime --------------------------------------------------------- */
/* selects the device and creates the library stream; idempotent */
int mugpu_init(int device);
const char *mugpu_last_error(void);
/* device name, CU count, HBM bytes, max LDS bytes per workgroup */
int mugpu_device_info(char *name, int name_len, int *cu_count,
                      size_t *hbm_bytes, int *lds_per_block);
/* field memory: in the reference every field is a std::vector<Real> owned
 * by a FieldCollection (src/common/field_typed.hh:443-449).  Blocks are
 * cached: mugpu_free keeps a block (up to a quarter of the device memory in
 * all) for the next mugpu_alloc of the same size, so repeated solves do not
 * map and unmap memory each time; the blocks are for work on the library
 * stream.  mugpu_free of a pointer mugpu_alloc did not return (or already
 * freed) fails with MUGPU_EARG.  mugpu_release_cached (no reference
 * counterpart) returns the cached blocks to the device after the library
 * stream's work; an allocation that fails does so by itself and retries. */
int mugpu_alloc(size_t bytes, void **d_ptr);
/* Test hook for the failure paths of the library's own device allocations
 * (no reference counterpart; for tests of out-of-memory handling): from now
 * on the allocations of cells, transform plans, projections, the
 * mixed-precision work vectors, serving resolution classes and the
 * bandwidth probes (not mugpu_alloc) succeed `after` more times, and
 * every later one fails with MUGPU_EMEMORY as if the device were full, until
 * the hook is called with after < 0 (the default: no injected failure). */
int mugpu_debug_fail_allocations(int after);
int mugpu_free(void *d_ptr);
int mugpu_release_cached(void);
int mugpu_upload(void *d_dst, const void *h_src, size_t bytes);
int mugpu_download(void *h_dst, const void *d_src, size_t bytes);
int mugpu_memset_zero(void *d_ptr, size_t bytes);
int mugpu_synchronize(void);
/* the library's stream (a hipStream_t), for callers that time kernels */
void *mugpu_stream(void);

/* ---- FFT engine ------------------------------------------------------
 * FFTEngineBase ctor (src/fft/fft_engine_base.hh:75-76,
 * fft_engine_base.cc:40-51) + FFTWEngine::initialise
 * (src/fft/fftw_engine.cc:52-117): batched r2c/c2r of nb_components
 * interleaved components (howmany=nc, stride nc). */
int mugpu_fft_create(int dim, const int *resolutions, int nb_components,
                     int precision, mugpu_fft *out);
int mugpu_fft_destroy(mugpu_fft plan);
/* FFTWEngine::fft (src/fft/fftw_engine.cc:130-143): unnormalised r2c,
 * d_real (N*nc reals) -> d_spectrum (Kpts*nc complex); input preserved */
int mugpu_fft_r2c(mugpu_fft plan, const void *d_real, void *d_spectrum);
/* FFTWEngine::ifft (src/fft/fftw_engine.cc:146-158): unnormalised c2r;
 * the spectrum is overwritten, as with the reference's FFTW c2r plan */
int mugpu_fft_c2r(mugpu_fft plan, void *d_spectrum, void *d_real);
/* FFTEngineBase::normalisation (src/fft/fft_engine_base.hh:145): 1/N */
double mugpu_fft_normalisation(mugpu_fft plan);
/* number of Hermitian k-points, N0*..*(N_{d-1}/2+1)
 * (src/common/ccoord_operations.hh:101-104) */
int64_t mugpu_fft_nb_fourier_points(mugpu_fft plan);

/* ---- projection ------------------------------------------------------
 * ProjectionSmallStrain / ProjectionFiniteStrainFast / ProjectionFiniteStrain
 * ctor + initialise (src/fft/projection_small_strain.cc:42-86,
 * projection_finite_strain_fast.cc:43-73, projection_finite_strain.cc:45-95).
 * The operator is generated from the wave-vector indices inside the kernel;
 * nothing is tabulated.  allow_even = 0 keeps the reference's odd-only
 * check (MUGPU_EPROJECTION); 1 zeroes the operator on every Nyquist index
 * of even axes.  The engine is owned by the projection afterwards
 * (src/fft/projection_base.cc:41-53) and destroyed with it. */
int mugpu_projection_create(mugpu_fft engine, int kind, const double *lengths,
                            int allow_even, mugpu_proj *out);
int mugpu_projection_destroy(mugpu_proj proj);
/* ProjectionBase::apply_projection (src/fft/projection_base.hh:120,
 * projection_default.cc:53-63, projection_finite_strain_fast.cc:77-87):
 * in place on d_field (N*nc reals, plan precision) */
int mugpu_projection_apply(mugpu_proj proj, void *d_field);
/* ProjectionBase::get_operator (src/fft/projection_base.hh:144): host copy
 * of the reference's stored operator, Kpts x (nc*nc) for the full-G
 * projections and Kpts x dim for the fast finite-strain one */
int mugpu_projection_get_operator(mugpu_proj proj, double *h_out);

/* ---- cell -------------------------------------------------------------
 * make_cell / CellBase (src/cell/cell_factory.hh:59-105,
 * src/cell/cell_base.cc:48-62): owns F ("Gradient"), P ("Piola-Kirchhoff-1"),
 * the tangent, the projection of the formulation (finite strain: fast
 * projection, small strain: small-strain projection), and materials.
 * precision selects the operator stage (tangent product, FFT, projection). */
int mugpu_cell_create(int dim, const int *resolutions, const double *lengths,
                      int formulation, int precision, int allow_even,
                      mugpu_cell *out);
int mugpu_cell_destroy(mugpu_cell cell);
int64_t mugpu_cell_nb_pixels(mugpu_cell cell);
int mugpu_cell_nb_components(mugpu_cell cell);
/* Material::make (src/materials/material_muSpectre_base.hh:228-237):
 * returns the material index in *mat.  young/poisson for LE1, LE2, HEP1;
 * tau_y0 and H for HEP1 (material_hyper_elasto_plastic1.cc:40-65) */
int mugpu_cell_add_material(mugpu_cell cell, int family, double young,
                            double poisson, double tau_y0, double H,
                            int *mat);
/* MaterialLinearElasticGeneric(name, C_voigt)
 * (src/materials/material_linear_elastic_generic.cc:40-65): C_voigt is the
 * vsize x vsize stiffness in Voigt notation, row-major (vsize 3 in 2D with
 * the order xx, yy, xy; 6 in 3D with xx, yy, zz, yz, xz, xy,
 * src/common/voigt_conversion.hh:118-125); C_ijkl = C_voigt(v(ij), v(kl)).
 * Green-Lagrange strain to PK2 stress, S = C : E (:164-176); the material
 * index is returned in *mat and pixels are added with
 * mugpu_material_add_pixels (no parameters). */
int mugpu_cell_add_material_generic(mugpu_cell cell, const double *C_voigt,
                                    int *mat);
/* MaterialBase::add_pixel and the LE2-LE4 overloads
 * (src/materials/material_base.cc:56-58, material_linear_elastic2.cc:57-64,
 * material_linear_elastic3.cc:57-68, material_linear_elastic4.cc:62-73):
 * pixels are linear indices; params per pixel: LE2 the eigenstrain
 * (dim*dim), LE3/LE4 (Young, Poisson); NULL otherwise */
int mugpu_material_add_pixels(mugpu_cell cell, int mat, int64_t count,
                              const int64_t *pixels, const double *params);
/* CellBase::initialise (src/cell/cell_base.cc:445-454): coverage check
 * (:486-526), material initialise, projection initialise */
int mugpu_cell_initialise(mugpu_cell cell);
int mugpu_cell_is_initialised(mugpu_cell cell);
/* device pointers of F and P (FP64, N*nc): the fields behind
 * Cell::get_strain_vector / get_stress_vector (src/cell/cell_base.hh:148) */
int mugpu_cell_strain(mugpu_cell cell, void **d_F);
int mugpu_cell_stress(mugpu_cell cell, void **d_P);
/* CellBase::set_uniform_strain (src/cell/cell_base.cc:102-106); h_strain is
 * dim*dim, column-major */
int mugpu_cell_set_uniform_strain(mugpu_cell cell, const double *h_strain);
/* F += h_increment on every pixel (src/solver/solvers.cc:142-145) */
int mugpu_cell_add_uniform_strain(mugpu_cell cell, const double *h_increment);
/* CellBase::evaluate_stress_tangent (src/cell/cell_base.cc:129-146):
 * every material on its own pixels; P and the tangent are updated */
int mugpu_cell_evaluate_stress_tangent(mugpu_cell cell);
/* CellBase::evaluate_stress (src/cell/cell_base.cc:112-124) */
int mugpu_cell_evaluate_stress(mugpu_cell cell);
/* copy of the tangent as the reference stores it, N x (nc*nc) FP64
 * (src/cell/cell_base.cc:323-334) */
int mugpu_cell_get_tangent(mugpu_cell cell, double *h_K);
/* CellBase::evaluate_projected_directional_stiffness
 * (src/cell/cell_base.cc:150-192): d_out = G (K : d_in); FP64 vectors */
int mugpu_cell_directional_stiffness(mugpu_cell cell, const void *d_in,
                                     void *d_out);
/* K : d_in only, no projection (src/cell/cell_base.cc:182-188) */
int mugpu_cell_tangent_product(mugpu_cell cell, const void *d_in,
                               void *d_out);
/* CellBase::apply_projection (src/cell/cell_base.cc:201-208), FP64 vector */
int mugpu_cell_apply_projection(mugpu_cell cell, void *d_vec);
/* CellBase::save_history_variables (src/cell/cell_base.cc:458-462) ->
 * StateField::cycle (src/common/statefield.hh:322-326) */
int mugpu_cell_save_history(mugpu_cell cell);
/* CellBase::get_globalised_internal_real_array (src/cell/cell_base.cc:
 * 371-441): "eps_p", "F_prev", "be" (HEP1, current values) or
 * "eigenstrain" (LE2); pixels without the field get zeros */
int mugpu_cell_get_internal(mugpu_cell cell, const char *name, double *h_out,
                            int *nb_components);

/* MaterialMuSpectre's internal state fields (material_muSpectre_base.hh,
 * StateField src/common/statefield.hh:206-326) of a HyperElastoPlastic1
 * material: one sub-field ("<prefix>, sub_field index i", i = 0, 1) of
 * "cumulated plastic flow εₚ" ("eps_p", 1 component), "Previous
 * placement gradient Fᵗ" ("F_prev", nc) or "Previous left Cauchy-Green
 * deformation bₑᵗ" ("be", nc), in the material's own pixel order
 * (the order of mugpu_material_add_pixels), into h_out (nb_pixels of the
 * material x nb_components; NULL to ask only for the shape).
 * *current_sub_field returns the sub-field that is current: 1 after
 * mugpu_cell_initialise (material_hyper_elasto_plastic1.cc:76-84), swapped by
 * each mugpu_cell_save_history (statefield.hh:318-326). */
int mugpu_material_get_state(mugpu_cell cell, int mat, const char *name,
                             int sub_field, double *h_out, int *nb_components,
                             int *current_sub_field);
/* the operator of the cell's projection (CellBase::get_projection,
 * src/cell/cell_base.hh:323, ProjectionBase::get_operator,
 * projection_base.hh:144): Kpts x (nc*nc) for small strain, Kpts x dim for
 * the fast finite-strain projection, as mugpu_projection_get_operator */
int mugpu_cell_projection_operator(mugpu_cell cell, double *h_out);

/* ---- vector kernels for the Newton driver (src/solver/solvers.cc:
 * 143-145,173-175,184-187; Communicator::sum is the identity on one GPU,
 * src/common/communicator.hh:122-124) ------------------------------------ */
/* sum of squares / dot over n doubles, deterministic order */
int mugpu_dot(const void *d_a, const void *d_b, int64_t n, double *h_out);
/* d_y = a*d_x + b*d_y over n doubles */
int mugpu_axpby(double a, const void *d_x, double b, void *d_y, int64_t n);
/* d_y = a*d_x */
int mugpu_scale_copy(double a, const void *d_x, void *d_y, int64_t n);

/* ---- SolverCG (src/solver/solver_cg.cc:54-110) -----------------------
 * Device-resident CG.  The solver owns x, r, p, Ap (solver_cg.cc:45-51);
 * the host loop keeps the reference's control flow and reads back r.r once
 * per iteration (mugpu_cg_step) or once per chunk of iterations
 * (mugpu_cg_steps).  A cell runs one CG solve at a time: its device
 * scalars carry the solve's r.r and beta from call to call.
 * mugpu_cg_start: x = 0; r = A x - b (= 0 - b: the operator of a zero
 *   field is zero exactly, so it is not applied; Ap is left unset);
 *   p = -r; returns rdr = r.r and rhs_norm2 = b.b.
 * mugpu_cg_step: [if update_p: p = -r + beta p with beta from the previous
 *   step] Ap = A p; alpha = rdr/(p.Ap); x += alpha p; r += alpha Ap;
 *   new_rdr = r.r; beta = new_rdr/rdr.  Returns new_rdr.
 * mugpu_cg_steps: up to max_steps iterations of mugpu_cg_step (update_p
 *   applies to the first; the later ones update p) queued at once, with
 *   the loop test of solver_cg.cc:70-96 applied on the device after each
 *   iteration: the next one runs while fewer than max_steps have run and
 *   new_rdr > tol2 (the first always runs).  *steps returns the iterations
 *   run (one whose r.r is NaN counts and stops the loop), *new_rdr the r.r
 *   of the last.  The host reads back once per call instead of once per
 *   iteration; the arithmetic is that of mugpu_cg_step.  Where the last
 *   axis has compiled passes (lengths 16, 32, 48, 64, 96, 128, 256, 512),
 *   the x update of an iteration is applied by the next iteration's first
 *   pass (which reads p anyway), so the x update of the last iteration of a call is left
 *   pending (its alpha kept by the cell): the next call with update_p = 1
 *   on the same x and p applies it, and mugpu_cg_finish applies it after
 *   the last call; mugpu_cg_step, mugpu_cg_start and a call on other
 *   vectors apply it first.  mugpu_free of that p applies it before the
 *   block is released; mugpu_free of that x, or destroying the cell, drops
 *   it.
 * mugpu_cg_finish: applies a pending x update (no-op otherwise).
 * The operator runs in FP64 (cells created with MUGPU_FP32 solve with
 * mugpu_cg_solve_mixed). */
int mugpu_cg_start(mugpu_cell cell, const void *d_rhs, void *d_x, void *d_r,
                   void *d_p, void *d_Ap, double *rdr, double *rhs_norm2);
int mugpu_cg_step(mugpu_cell cell, int update_p, void *d_x, void *d_r,
                  void *d_p, void *d_Ap, double *new_rdr);
int mugpu_cg_steps(mugpu_cell cell, int update_p, void *d_x, void *d_r,
                   void *d_p, void *d_Ap, double tol2, int max_steps,
                   int *steps, double *new_rdr);
int mugpu_cg_finish(mugpu_cell cell);
/* mugpu_cg_steps on cells of at most 2^20 values (nb_pixels x nb_components)
 * replays each chunk from a captured hipGraph (kept on the cell per set of
 * vectors, chunk length and form of the first iteration; the arithmetic is
 * the same), because there the host's kernel launches, not the device, set
 * the pace.  Not while kernels are timed one by one (mugpu_timing_start).
 * enable = 0 turns this off for the process, 1 (the default) on.  No
 * reference counterpart. */
int mugpu_cg_graphs(int enable);

/* ---- mixed-precision CG with FP64 refinement (no reference counterpart;
 * it replaces SolverCG::solve for cells created with MUGPU_FP32) --------
 * Inner CG on FP32 vectors with the FP32 operator and FP64 dot products
 * and alpha/beta; the outer loop recomputes the true residual b - A x with
 * the FP64 operator and corrects x in FP64.  Stops when
 * |b - A x|^2 <= tol^2 |b|^2 or after maxiter inner iterations in total.
 * Each inner solve stops at inner_tol relative, or earlier once its
 * residual, scaled back to b - A x, is tol |b| / 2 (the outer test passes
 * with a factor of two to spare), so that the last inner solve of a solve
 * whose tol is above inner_tol^2 goes no further than tol asks.
 * *iterations returns the inner CG iterations, *refinements the outer
 * FP64 residual evaluations.  Returns MUGPU_ECONVERGENCE when not
 * converged. */
int mugpu_cg_solve_mixed(mugpu_cell cell, const void *d_rhs, void *d_x,
                         double tol, int maxiter, double inner_tol,
                         int64_t *iterations, int *refinements,
                         double *rel_residual);

/* ---- mixed-precision CG without device allocation (no reference
 * counterpart) -------------------------------------------------------------
 * For callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc): mugpu_cg_solve_mixed with the same
 * arguments, results and errors, whose work vectors (four FP32 and one FP64
 * vector of nb_pixels x nb_components values, plus the loop's control words)
 * live in a fixed device array of MUGPU_NOALLOC_SCRATCH_BYTES instead of
 * memory allocated on the cell.  The array is part of the library, one per
 * GPU, held until the process ends; the call allocates nothing (the cell,
 * the right-hand side and x are the caller's).  Cells with more than
 * MUGPU_NOALLOC_MAX_VALUES values (e.g. 3D up to 64^3, 2D up to 512^2) are
 * rejected with MUGPU_EARG.  Every call on a GPU uses the same array: calls
 * that overlap on one GPU (from several threads) race; calls in sequence do
 * not depend on each other. */
#define MUGPU_NOALLOC_SCRATCH_BYTES (64ll << 20)
#define MUGPU_NOALLOC_MAX_VALUES 2796000ll
int mugpu_cg_solve_mixed_noalloc(mugpu_cell cell, const void *d_rhs, void *d_x,
                                 double tol, int maxiter, double inner_tol,
                                 int64_t *iterations, int *refinements,
                                 double *rel_residual);

/* ---- serving loop (no reference counterpart) ----------------------------
 * Many independent RVE solves of uneven resolution on one GPU.  Each request
 * is one small-strain load step of newton_cg (src/solver/solvers.cc:44-217)
 * on a cell with per-voxel isotropic linear elasticity (Young's modulus and
 * Poisson ratio per voxel, MaterialLinearElastic4, material_linear_elastic4
 * .hh:189-210), solved with the mixed-precision SolverCG (FP32 inner CG,
 * FP64 refinement).  Requests are admitted first come, first served while
 * their device footprint fits the HBM budget; the inner CG iterations of all
 * running requests of one resolution run in shared, graph-captured batched
 * launches; a request's Newton and refinement steps run between them on a
 * second stream.  Each request computes exactly what it computes alone.
 * Resolutions: every axis in {16, 32, 48, 64, 96, 128, 256, 512}; even
 * lengths use the Nyquist policy (allow_even). */
typedef struct mugpu_server_s *mugpu_server;

typedef struct {
  size_t hbm_budget;    /* device bytes for admitted requests; 0: 3/4 of free */
  int max_batch;        /* slots per resolution (1..64) */
  int iters_per_launch; /* inner CG iterations per batched launch (>= 1) */
  int use_graphs;       /* capture the batched iterations in hipGraphs */
  int keep_fields;      /* keep strain and stress fields of finished requests */
  double cg_tol;        /* SolverCG tolerance, relative to |rhs| */
  int cg_maxiter;       /* SolverCG maxiter (also bounds the Newton loop) */
  double inner_tol;     /* FP32 inner CG tolerance */
  double newton_tol;
  double equil_tol;
} mugpu_server_config;

int mugpu_server_create(const mugpu_server_config *config, mugpu_server *out);
int mugpu_server_destroy(mugpu_server server);
/* device bytes one request of this resolution occupies while it runs */
int mugpu_server_footprint(int dim, const int *resolutions, size_t *bytes);
/* queue one solve; young/poisson per voxel (row-major pixels), strain the
 * symmetric dim x dim macroscopic strain (column-major); cg_maxiter 0 takes
 * the server's.  The per-voxel parameters are copied (as Lame pairs) before
 * the call returns; for requests of 131,072 voxels or more the copy runs on
 * up to eight short-lived host threads.  MUGPU_ESOLVER for an asymmetric
 * strain, MUGPU_EMEMORY when the footprint exceeds the whole budget or the
 * host copy cannot be allocated, MUGPU_EARG for a resolution without
 * compiled passes. */
int mugpu_server_submit(mugpu_server server, int dim, const int *resolutions,
                        const double *lengths, const double *young,
                        const double *poisson, const double *strain,
                        int cg_maxiter, int64_t *id);
/* one turn of the loop: admit, advance every request and batch whose device
 * work has completed, launch what is ready; with wait != 0 and nothing
 * completed, block until some device work completes.  *unfinished counts
 * queued and running requests. */
int mugpu_server_progress(mugpu_server server, int wait, int *unfinished);
/* ids of the requests finished since the last call (at most capacity) */
int mugpu_server_poll(mugpu_server server, int64_t *ids, int capacity,
                      int *count);
/* result of a finished request: status (MUGPU_OK, or MUGPU_ECONVERGENCE when
 * the CG failed), whether newton_cg converged, Newton and CG iteration
 * counts, the homogenised stress (dim*dim, column-major) and, when the
 * server keeps fields and the pointers are not NULL, the strain and stress
 * fields (nb_pixels x dim*dim, AoS) */
int mugpu_server_result(mugpu_server server, int64_t id, int *status,
                        int *converged, int *newton_iterations,
                        int64_t *cg_iterations, double *mean_stress,
                        double *strain, double *stress);
int mugpu_server_release(mugpu_server server, int64_t id);
/* counters: [0] submitted, [1] admitted, [2] finished, [3] failed,
 * [4] batched launches, [5] summed slots over batched launches,
 * [6] peak device bytes in use, [7] budget, [8] peak requests running */
int mugpu_server_stats(mugpu_server server, int64_t *out, int n);

/* ---- instrumentation (no reference counterpart) ------------------------
 * HIP-event timing of one kernel class (or of all, MUGPU_K_ALL) on the
 * library stream, used by the benchmark to measure kernel durations inside
 * a timed run.  mugpu_timing_stop returns the launches and milliseconds of
 * the selected class(es); mugpu_timing_class then reports one class of the
 * run that was just stopped. */
#define MUGPU_K_LAST_R2C 0  /* last-axis r2c pass (with the K:p stage)      */
#define MUGPU_K_MIDDLE 1    /* middle-axis c2c passes (3D)                  */
#define MUGPU_K_AXIS0 2     /* first-axis pass with the projection          */
#define MUGPU_K_LAST_C2R 3  /* last-axis c2r pass (with the p.Ap stage)     */
#define MUGPU_K_MATERIAL 4  /* constitutive kernels                         */
#define MUGPU_K_CG_UPDATE 5 /* x, r update with the r.r partial sums        */
#define MUGPU_K_ALL 6       /* every class above                             */
int mugpu_timing_start(int kernel_class);
int mugpu_timing_stop(int64_t *launches, double *total_ms);
int mugpu_timing_class(int kernel_class, int64_t *launches, double *total_ms);
/* autotuner measurement: average duration in ms of one compile-time-length
 * FFT pass (MUGPU_PASS_*) in one variant (bit 0: 8 values per thread instead
 * of 16 or 12; bit 1, for MUGPU_PASS_R2C, MUGPU_PASS_C2R and
 * MUGPU_PASS_AXIS0: half the rows or k-points per workgroup; MUGPU_EARG for
 * a variant that does not exist at this length)
 * over `reps` launches, each timed alone after a copy that evicts the
 * last-level cache, on zeroed scratch buffers of
 * the grid `resolutions`; iso selects the isotropic tangent for the r2c
 * pass.  The chosen variants are compiled into the library
 * (muspectre_amd/csrc/fft_tuned.inc, written by muspectre_amd/autotune.py). */
#define MUGPU_PASS_R2C 0          /* last-axis r2c with p update and K:p      */
#define MUGPU_PASS_MIDDLE 1       /* middle-axis c2c (3D)                     */
#define MUGPU_PASS_AXIS0 2        /* first axis with the projection           */
#define MUGPU_PASS_C2R 3          /* last-axis c2r with the p.Ap partials     */
#define MUGPU_PASS_AXIS0_PLAIN 4  /* first-axis c2c of the engine transforms  */
int mugpu_fft_tune_pass(int dim, const int *resolutions, int precision, int pass,
                        int variant, int iso, int reps, double *avg_ms);
/* Test hook (no reference counterpart; for tests that run every variant a
 * tuned table could choose): plans created from now on use `variant` of
 * `pass` (MUGPU_PASS_*) in `precision` on every axis length where that
 * variant exists, in place of the tuned one; variant < 0 restores the table. */
int mugpu_debug_force_variant(int pass, int precision, int variant);
/* streaming device copy of `bytes` (multiple of 16), the bandwidth
 * reference for the roofline and for the HBM counter calibration */
int mugpu_copy(void *d_dst, const void *d_src, size_t bytes);
/* achievable device bandwidth: GB/s of six streams over `bytes` in two
 * fresh device blocks, each timed with HIP events over `reps` launches after
 * one untimed launch: copies (bytes read + written) gbs[0] mugpu_copy's
 * kernel, gbs[1] a streaming copy with four 16-byte loads in flight per lane,
 * gbs[2] the same with non-temporal loads and stores, gbs[3]
 * hipMemcpyAsync device to device; gbs[4] a read-only stream (non-temporal
 * loads, four in flight per lane), gbs[5] a write-only stream (non-temporal
 * stores).  gbs holds 6 values. */
int mugpu_copy_bandwidth(size_t bytes, int reps, double *gbs);
/* bandwidth of the first-axis pass's HBM pattern alone (no reference
 * counterpart; a roofline anchor for the fused first-axis kernel): a fresh
 * device block of n rows (n a multiple of 16) of plane k-points with
 * nb_components complex values each (precision MUGPU_FP64 or MUGPU_FP32) is
 * read and written in place by tiles of kpts_per_tile k-points, every thread
 * holding 16 values of one column, in the tile order of that kernel; *gbs =
 * bytes read + written per second over `reps` launches after one untimed
 * launch.  kpts_per_tile * nb_components * n / 16 <= 1024. */
int mugpu_line_bandwidth(int precision, int n, int64_t plane, int nb_components,
                         int kpts_per_tile, int reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif
