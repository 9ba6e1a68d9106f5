// Host-side structures and kernel launchers shared by the HIP translation
// units of libmugpu.so.  Nothing here is part of the public C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "fft_core.hpp"
#include "mugpu.h"

namespace mugpu {

// ---- error handling --------------------------------------------------------
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
#define MUGPU_HIP(call)                                                   \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return ::mugpu::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)
#define MUGPU_TRY(call)          \
  do {                           \
    int rc_ = (call);            \
    if (rc_ != MUGPU_OK) return rc_; \
  } while (0)

// Device allocation of the library's own buffers (cells, tangents, plans,
// work vectors): a failed hipMalloc has its sticky error cleared (the next
// launch check would report it otherwise), the field cache is handed back to
// the device after the library stream's work, and the allocation is retried
// once.  On failure *p is nullptr and MUGPU_EMEMORY is returned with `what`
// in the message.
int dev_alloc(void** p, size_t bytes, const char* what);
template <typename T>
inline int dev_alloc(T** p, size_t bytes, const char* what) {
  return dev_alloc(reinterpret_cast<void**>(p), bytes, what);
}

// Squared stopping threshold of one inner FP32 CG solve of the mixed-precision
// refinement, for the inner system A d = r64 / |r64| whose start residual has
// squared norm rdr: the inner solve stops at inner_tol relative or, when that
// is more than the outer tolerance still asks for, once |r64| times its
// relative residual is half the outer threshold (tol2 = tol^2 |b|^2, rr =
// |r64|^2).  FP32 CG converges evenly to about 1e-4 relative and slows beyond
// it, so the first inner solve stops at inner_tol; the last one of a solve at
// a tolerance above inner_tol^2 stops as soon as the outer test will pass.
inline double inner_threshold2(double inner_tol, double tol2, double rr, double rdr) {
  const double need2 = 0.25 * tol2 / rr;
  const double rel2 = inner_tol * inner_tol;
  return (need2 > rel2 ? need2 : rel2) * rdr;
}

hipStream_t stream();
// route the library's launches to another stream (nullptr: the library
// stream); used by the serving loop, which drives two streams from one thread
void set_current_stream(hipStream_t s);
// while a stream is captured into a graph, the per-launch event timing is off
void set_capturing(bool on);
int max_lds_per_block();
// per-launch event timing of one kernel class (MUGPU_K_*), off by default
void timing_begin(int kclass);
void timing_end(int kclass);
// position in the timing record, and dropping the launches recorded since a
// position (launches that returned at once; not counted anywhere)
int64_t timing_mark();
void timing_void(int64_t from);

// ---- tangent representations of the directional-stiffness product --------
// FULL: the reference's stored K, nc*nc entries per voxel, entry r + nc*c is
//       K(r, c) (src/common/T4_map_proxy.hh:83-105)
// ISO : (lambda, mu) per voxel; dP = lambda tr(dF) I + mu (dF + dF^T), the
//       contraction with C = lambda Itrac + 2 mu Isymm
// HEP : dim^3 per voxel, M(i,m,n) at i + d m + d^2 n, for the j-independent
//       Kirchhoff->PK1 tangent K(i,m,j,n) = M(i,m,n)
//       (src/materials/materials_toolbox.hh:394-406)
// SYM : the upper triangle r <= c of a major-symmetric K (finite-strain
//       hyperelastic laws and the consistent elasto-plastic tangent),
//       nc (nc + 1) / 2 entries, row by row
//
// Storage is voxel-blocked: blocks of kTanBlock consecutive voxels, entry
// major inside a block, so that lanes on consecutive voxels read one
// contiguous run per entry (coalesced) while a voxel's entries stay in one
// block.
enum TangentKind { TAN_FULL = 0, TAN_ISO = 1, TAN_HEP = 2, TAN_SYM = 3 };
constexpr int kTanBlock = 64;
__host__ __device__ inline int tangent_stride(int kind, int dim) {
  const int nc = dim * dim;
  return kind == TAN_FULL ? nc * nc
                          : kind == TAN_ISO ? 2 : kind == TAN_HEP ? dim * dim * dim : nc * (nc + 1) / 2;
}
// number of values allocated for a tangent over npix voxels
inline int64_t tangent_values(int kind, int dim, int64_t npix) {
  return (npix + kTanBlock - 1) / kTanBlock * kTanBlock * tangent_stride(kind, dim);
}
// offset of entry q of voxel v for a stride of S entries per voxel
__host__ __device__ inline int64_t tan_off(int64_t v, int q, int S) {
  return (v / kTanBlock) * ((int64_t)S * kTanBlock) + (int64_t)q * kTanBlock + (v % kTanBlock);
}
// packed index of K(r, c), r <= c, in the SYM representation of an n x n matrix
__host__ __device__ constexpr int sym_index(int r, int c, int n) {
  return r * n - r * (r - 1) / 2 + (c - r);
}

// device scalar block used by the CG kernels
// SC_TOL2: squared stopping threshold of a batched inner CG
// SC_XALPHA: alpha of a pending x update (kept apart from SC_ALPHA, which
// the next solve on the cell overwrites)
enum Scalar { SC_RDR = 0, SC_PAP, SC_ALPHA, SC_NEWRDR, SC_BETA, SC_TMP0, SC_TMP1, SC_TOL2,
              SC_XALPHA, SC_COUNT };

// ---- FFT plan -----------------------------------------------------------------
struct FftPlan {
  int dim = 0, nc = 0, prec = MUGPU_FP64;
  int res[3] = {1, 1, 1}, hres[3] = {1, 1, 1};
  int64_t N = 0, kpts = 0;
  AxisPlan ax[3];
  void* tw64[3] = {nullptr, nullptr, nullptr};  // double2 tables W_n^m
  void* tw32[3] = {nullptr, nullptr, nullptr};  // float2 tables
  // z-pass: rows per workgroup; strided passes: lines per workgroup
  int zrb = 1;
  int yW = 1;
  int xW = 1;      // plain axis-0 pass
  int xWk = 1;     // fused axis-0 projection pass: k-points per workgroup
  int xNT = 256;   // threads of the fused pass
  void* partials = nullptr;  // double per z-pass workgroup
  int64_t nzblocks = 0;      // z-pass workgroups, the largest of nzb
  int64_t nzb[2] = {0, 0};   // c2r workgroups (p.Ap partials) per precision (0 FP64, 1 FP32)
  bool fastax[3] = {false, false, false};  // compile-time-length passes (fft_fast.hip)
  // variant of each compile-time-length pass (bit 0: 8 values per thread
  // instead of 16 or 12; bit 1: half the rows or k-points per workgroup, for
  // the last-axis and first-axis projection passes), per precision (0 FP64,
  // 1 FP32), per pass and, for the r2c pass, per tangent (0 full matrices,
  // 1 isotropic)
  unsigned char fvar[2][5][2] = {};
  // while set: every pass kernel returns at once when active[0] == 0 (CG
  // iterations launched ahead of the host's stopping test)
  const int* active = nullptr;
};
// passes with variants
enum FastVariantPass { FV_R2C = 0, FV_STRIDED1 = 1, FV_AXIS0 = 2, FV_C2R = 3, FV_STRIDED0 = 4, FV_COUNT = 5 };
// variant of a pass for an axis length, dimension, precision and tangent
// class, from the tuned table (fft_tuned.inc)
int fast_variant(int pass, int n, int dim, int prec, int iso);
// make fast_variant return `variant` for a pass and precision wherever that
// variant exists (-1: the tuned table again); plans read it when created
void force_variant(int pass, int prec, int variant);
// whether a pass variant exists (and differs from variant 0's tile) for
// this length, dimension and precision
bool fast_variant_valid(int pass, int n, int dim, int prec, int variant);

int plan_init(FftPlan& p, int dim, const int* res, int nc, int prec);
void plan_free(FftPlan& p);

// projection parameters handed to the fused axis-0 kernel
struct ProjParams {
  int kind = -1;  // MUGPU_PROJ_*; -1 none
  int allow_even = 0;
  double L[3] = {1, 1, 1};
  double norm = 1.0;
};

// input stage of the last-axis r2c pass
// IN_TAN_PUPDATE_X: kernel form of IN_TAN_PUPDATE that also applies x +=
// alpha p_old (chosen by the launcher when OpSpec::xvec is set)
enum InMode { IN_PLAIN = 0, IN_TAN = 1, IN_TAN_PUPDATE = 2, IN_TAN_PUPDATE_X = 3 };
// output stage of the last-axis c2r pass
enum OutMode { OUT_PLAIN = 0, OUT_DOT = 1 };

struct OpSpec {
  int in_mode = IN_PLAIN;
  const void* in = nullptr;      // real field (plan precision)
  const void* rvec = nullptr;    // CG residual for IN_TAN_PUPDATE
  void* pout = nullptr;          // updated p for IN_TAN_PUPDATE
  // IN_TAN_PUPDATE with compile-time-length last axis only: x += alpha p_old
  // on the way in (alpha = scal[SC_ALPHA], the previous iteration's), so
  // that the CG update need not read p again
  void* xvec = nullptr;
  const double* scal = nullptr;  // device scalars (SC_BETA)
  int tan_kind = TAN_FULL;
  const void* tangent = nullptr; // plan precision
  int out_mode = OUT_PLAIN;
  void* out = nullptr;           // real field (plan precision)
  const void* dotv = nullptr;    // vector dotted with the output for OUT_DOT
};

// full operator: r2c (with input stage) -> [axis passes] -> projection ->
// [inverse axis passes] -> c2r (with output stage); spectrum is workspace
int run_operator(FftPlan& p, int prec, const OpSpec& op, const ProjParams& pj,
                 void* spectrum);
// engine transforms (no projection)
int run_r2c(FftPlan& p, int prec, const void* in, void* spectrum);
int run_c2r(FftPlan& p, int prec, void* spectrum, void* out);

struct BatchSlot;
// passes for axis lengths 16, 32, 48, 64, 96, 128, 256, 512 (fft_fast.hip)
bool fast_size(int n);
// rows per workgroup of the last-axis passes in a variant
int fast_rows_per_block(int n, int dim, int variant);
int fast_r2c(FftPlan& p, int prec, const OpSpec& op, void* spec);
int fast_c2r(FftPlan& p, int prec, const OpSpec& op, void* spec);
int fast_strided(FftPlan& p, int prec, int ax, void* spec, int sign);
int fast_axis0_project(FftPlan& p, int prec, const ProjParams& pj, void* spec);
// the HBM pattern of the first-axis pass alone (n rows of plane k-points of nc
// complex components, tiles of kp k-points), in place (fft_fast.hip)
int line_pattern(void* data, int prec, int n, int64_t plane, int nc, int kp);
// batched FP32 operator G:(K:p) with the p update and the p . Ap partials,
// one RVE per slot (fft_fast.hip)
int fast_batch_operator(FftPlan& p, const BatchSlot* d_slots, int nslots);
// batched k_finalize (alpha or beta + stopping test) and x, r update (cg.hip)
int batch_finalize(const BatchSlot* d_slots, int nslots, int64_t count, int mode);
int batch_cg_update(const BatchSlot* d_slots, int nslots, int64_t n);

// ---- batched inner CG (serving) -------------------------------------------
// One RVE of a batch of equal-resolution RVEs whose FP32 inner CG iterations
// run in the same launches (slot = blockIdx.z).  ctl[0] active flag, ctl[1]
// iterations done, ctl[2] iteration limit; a finished slot is skipped by
// every later launch, so a solve stops exactly where it would stop alone.
struct BatchSlot {
  const float* tan;
  float* x;
  float* r;
  float* p;
  float* Ap;
  float2* spec;
  double* scal;   // SC_* block of this RVE
  double* zpart;  // per-workgroup partials of the c2r pass (p . Ap)
  double* upart;  // per-workgroup partials of the update (r . r)
  int* ctl;
  int tan_kind;
  int kind;        // projection (MUGPU_PROJ_*)
  int allow_even;
  double L[3];
  double norm;
};
enum BatchCtl { CTL_ACTIVE = 0, CTL_ITER = 1, CTL_MAXIT = 2, CTL_COUNT = 4 };

// one inner CG iteration for the slots [0, nslots) of a batch (FP32 operator,
// FP64 scalars); slots must share the plan's resolution
int batch_cg_iteration(FftPlan& p, const BatchSlot* d_slots, int nslots);

// ---- reductions and vector kernels (cg.hip) --------------------------------
int dot_device(int prec, const void* a, const void* b, int64_t n, double* d_partials,
               double* d_scal, int slot);
// ctl (optional, CTL_* words): skip when ctl[CTL_ACTIVE] == 0; FIN_BETA then
// also applies SolverCG's loop test (ctl[CTL_ITER] + 1 < ctl[CTL_MAXIT] and
// r.r > scal[SC_TOL2]) and clears CTL_ACTIVE when it fails
int finalize(const double* d_partials, int64_t count, double* d_scal, int mode,
             int* ctl = nullptr);
// ctl = {active, 0 iterations, maxit}, scal[SC_TOL2] = tol2 (on the stream)
int set_ctl(int* ctl, int maxit, double* d_scal, double tol2);
enum FinalizeMode { FIN_ALPHA = 0, FIN_BETA = 1, FIN_STORE_TMP0 = 2, FIN_STORE_TMP1 = 3,
                    FIN_STORE_RDR = 4 };
// x == nullptr: x is updated elsewhere (OpSpec::xvec of the next r2c)
int cg_update(int prec, void* x, void* r, const void* p, const void* Ap, const double* d_scal,
              int64_t n, double* d_partials, int* ctl = nullptr);
// x += alpha p with alpha = scal[slot] (the update left over when the CG
// x update rides on the next iteration's r2c)
int axpy_alpha(int prec, void* x, const void* p, const double* d_scal, int64_t n,
               int slot = SC_ALPHA);
// scal[to] = scal[from] on the device
int scal_copy(double* d_scal, int from, int to);
int cg_init_vectors(int prec, const void* Ax, const void* b, void* r, void* p, int64_t n);
int axpby(int prec, double a, const void* x, double b, void* y, int64_t n);
int convert(int from_prec, const void* src, int to_prec, void* dst, int64_t n, double scale);
// x64 += scale d32; with p32: scale (d32 + scal[SC_ALPHA] p32)
int add_scaled_convert(void* x64, const void* d32, double scale, int64_t n,
                       const void* p32 = nullptr, const double* d_scal = nullptr);
// inner-solve start of the mixed CG: r32 = -s r64, p32 = -r32, x32 = 0,
// r32.r32 finalized into scal (mode as for finalize)
int mixed_start(const double* r64, double s, float* r32, float* p32, float* x32, int64_t n,
                double* d_partials, double* d_scal, int mode);
// r = b - r (r holds A x on entry), r.r finalized into scal
int residual_dot(const double* b, double* r, int64_t n, double* d_partials, double* d_scal,
                 int mode);
int fill_uniform(double* F, const double* h_val, int nc, int64_t npix, int add);
constexpr int kRedBlocks = 4096;  // capacity of the partial-sum buffers
constexpr int kRedGrid = 1024;    // default workgroup cap of the grid-stride kernels

// ---- materials (materials.hip) ----------------------------------------------
struct MaterialDev {
  int family = 0;
  double young = 0, poisson = 0, lambda = 0, mu = 0, bulk = 0, tau_y0 = 0, H = 0;
  std::vector<int64_t> h_pix;
  std::vector<double> h_params;  // per pixel
  int ppn = 0;
  std::vector<double> h_C;       // MUGPU_LEG: C_ijkl at (i + d j) + d^2 (k + d l)
  double* d_C = nullptr;
  bool major_sym = true;         // MUGPU_LEG: C_ijkl == C_klij
  int* d_pix = nullptr;
  double* d_params = nullptr;
  double* Fp[2] = {nullptr, nullptr};
  double* be[2] = {nullptr, nullptr};
  double* ep[2] = {nullptr, nullptr};
  int cur = 0;
  int64_t n_pixels = -1;  // pixel count when h_pix is not kept on the host
};

int eval_material(const MaterialDev& m, int dim, int form, const double* F, double* P,
                  int tan_kind, double* tangent, int need_tangent);

// serving: the material stage of several requests of one resolution in one
// launch (one request per blockIdx.y): small-strain per-voxel isotropic
// elasticity (MaterialLinearElastic4 over all pixels), writing the stress P,
// the right-hand side -P, the TAN_ISO tangent and its FP32 copy, bit for bit
// what eval_material and convert write
struct ServeMatSlot {
  const double* F;
  const double* params;  // (lambda, mu) per voxel
  double* P;
  double* rhs;
  double* tan64;
  float* tan32;
};
constexpr int kServeMatSlots = 64;  // requests per launch (passed by value)
int serve_material_batch(int dim, const ServeMatSlot* slots, int nslots, int64_t n);

}  // namespace mugpu
