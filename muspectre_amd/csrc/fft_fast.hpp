// FFT passes for axis lengths known at compile time (16, 32, 48, 64, 96,
// 128, 256, 512: the grids of the large and of the serving workloads):
// building blocks, workgroup shapes and kernels.  The launchers and the
// dispatch on length and variant are in fft_fast_r2c.hip, fft_fast_c2r.hip,
// fft_fast_strided.hip, fft_fast_axis0.hip and fft_fast_batch.hip, each
// instantiating its own kernels; fft_fast.hip holds the tuned variant table.
// Same stages and fusions as the runtime-length passes in fft_passes.hip,
// laid out for CDNA:
//
//  * lines across lanes: a workgroup owns L lines (columns of a tile) and
//    P = n/E threads per line; thread (l, p) holds E complex values of line
//    l in registers (E = 16, or 12 for the lengths with a factor 3, or 8).
//    Consecutive lanes work on consecutive lines, so every LDS access of a
//    pass is unit stride across lanes (element m of line l at s[m * LS + l],
//    LS = L or L + 1, whichever is odd), and on the strided axes every HBM
//    access is a contiguous run of L values;
//  * register-resident Stockham passes of radix 16/8/4/3/2 with every index
//    a compile-time shift or mask (the span before the radix-3 pass is a
//    power of two; no integer division), one LDS exchange between passes;
//  * on the strided axes the first pass reads HBM straight into registers
//    and the last pass writes HBM straight from registers.
//
// Stockham pass of radix R after a span NS (product of the earlier radices),
// butterfly j < n/R, k = j mod NS:
//   v_t = x[j + t n/R] W_{NS R}^{t k},  v = DFT_R(v),  y[(j/NS) NS R + k + t NS] = v_t
#pragma once
#include <type_traits>

#include "mugpu_internal.hpp"
#include "op_stages.hpp"

namespace mugpu {

namespace {

// complex values per thread: E = 16 (12 for the lengths with a factor 3), or
// E = 8 with twice the threads per line (n / E) on the same tile; ft(n) is
// the threads per line of the default plan, which fixes the tile shapes
__host__ __device__ constexpr int fe_default(int n) { return n % 3 == 0 ? 12 : 16; }
__host__ __device__ constexpr int ft(int n) { return n / fe_default(n); }

// radices of the passes of length n with E values per thread: each divides
// E, and the span before every pass is a power of two
__host__ __device__ constexpr int fast_npass(int n, int e) {
  return e == 8 ? (n == 16 || n == 32 || n == 64 ? 2 : 3)
                : (n == 16 ? 1 : n == 512 || n == 48 ? 3 : n == 96 ? 4 : 2);
}
__host__ __device__ constexpr int fast_radix(int n, int e, int i) {
  return e == 8 ? (n == 16    ? (i == 0 ? 8 : 2)
                   : n == 32  ? (i == 0 ? 8 : 4)
                   : n == 64  ? 8
                   : n == 128 ? (i < 2 ? 8 : 2)
                   : n == 256 ? (i < 2 ? 8 : 4)
                              : 8)
                : (n == 16    ? 16
                   : n == 32  ? (i == 0 ? 8 : 4)
                   : n == 48  ? (i < 2 ? 4 : 3)
                   : n == 64  ? 8
                   : n == 96  ? (i < 2 ? 4 : i == 2 ? 2 : 3)
                   : n == 128 ? (i == 0 ? 16 : 8)
                   : n == 256 ? 16
                              : 8);
}

template <int N, int E, int R>
__device__ __forceinline__ int in_idx(int p, int b, int t) {
  return p + (N / E) * b + t * (N / R);
}
template <int N, int E, int R, int NS>
__device__ __forceinline__ int out_idx(int p, int b, int t) {
  const int j = p + (N / E) * b;
  return (j / NS) * NS * R + (j & (NS - 1)) + t * NS;
}

// twiddles and butterflies of pass (R, NS) on this thread's E values
template <typename T, int N, int E, int R, int NS, int SG>
__device__ __forceinline__ void bfly_pass(c_t<T>* v, int p, const c_t<T>* __restrict__ tw) {
  constexpr int BPT = E / R;
  static_assert(BPT * R == E && (NS & (NS - 1)) == 0, "radix plan");
  constexpr int TWS = N / (NS * R);
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    if (NS > 1) {
      const int k = (p + (N / E) * b) & (NS - 1);
#pragma unroll
      for (int t = 1; t < R; ++t) {
        c_t<T> w = tw[t * k * TWS];
        if (SG > 0) w.y = -w.y;
        v[b * R + t] = cmul(v[b * R + t], w);
      }
    }
    Butterfly<T, R>::run(v + b * R, SG, nullptr, 0);
  }
}

// Runs pass I (inputs in v) and the passes after it; between passes the
// values go through the workgroup's LDS lines (at(m) is element m of this
// thread's line).  On return v holds the outputs of the last pass, to be
// stored at out_idx<N, R_last, N / R_last>.
template <typename T, int N, int E, int SG, int I, int NS>
struct Passes {
  static constexpr int R = fast_radix(N, E, I);
  template <class At>
  __device__ static void run(c_t<T>* v, int p, const c_t<T>* __restrict__ tw, At at) {
    bfly_pass<T, N, E, R, NS, SG>(v, p, tw);
    if constexpr (I + 1 < fast_npass(N, E)) {
      constexpr int R2 = fast_radix(N, E, I + 1);
#pragma unroll
      for (int b = 0; b < E / R; ++b)
#pragma unroll
        for (int t = 0; t < R; ++t) at(out_idx<N, E, R, NS>(p, b, t)) = v[b * R + t];
      __syncthreads();
#pragma unroll
      for (int b = 0; b < E / R2; ++b)
#pragma unroll
        for (int t = 0; t < R2; ++t) v[b * R2 + t] = at(in_idx<N, E, R2>(p, b, t));
      __syncthreads();
      Passes<T, N, E, SG, I + 1, NS * R>::run(v, p, tw, at);
    }
  }
};

template <int N, int E>
struct Last {
  static constexpr int R = fast_radix(N, E, fast_npass(N, E) - 1);
  static constexpr int NS = N / R;
};

// load the inputs of the first pass: v[b R0 + t] = ld(in index)
template <typename T, int N, int E, class Ld>
__device__ __forceinline__ void load_first(c_t<T>* v, int p, Ld ld) {
  constexpr int R0 = fast_radix(N, E, 0);
#pragma unroll
  for (int b = 0; b < E / R0; ++b)
#pragma unroll
    for (int t = 0; t < R0; ++t) v[b * R0 + t] = ld(in_idx<N, E, R0>(p, b, t));
}
// store the outputs of the last pass: st(out index, value)
template <typename T, int N, int E, class St>
__device__ __forceinline__ void store_last(const c_t<T>* v, int p, St st) {
  constexpr int R = Last<N, E>::R, NS = Last<N, E>::NS;
#pragma unroll
  for (int b = 0; b < E / R; ++b)
#pragma unroll
    for (int t = 0; t < R; ++t) st(out_idx<N, E, R, NS>(p, b, t), v[b * R + t]);
}

template <typename T>
__device__ __forceinline__ c_t<T> czero() {
  c_t<T> z;
  z.x = T(0);
  z.y = T(0);
  return z;
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b runs on
// XCD b mod 8), each XCD with its own L2.  Tiles next to each other in HBM
// share cache lines at their edges, so the tile of workgroup b is chosen to
// give every XCD one contiguous range of tiles: neighbours then meet in the
// same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t total) {
  const int64_t q = total / kXcds, r = total - q * kXcds;
  const int64_t x = b % kXcds, local = b / kXcds;
  return x * q + (x < r ? x : r) + local;
}

// ---- shapes of the workgroups -----------------------------------------------
// last-axis passes: rows per workgroup (even: two rows share a complex
// line); 288 threads in 3D, 256 in 2D; h = 1: half the rows (half the LDS,
// more workgroups per CU) where that leaves an even count
__host__ __device__ constexpr int z_rows0(int n, int dim) {
  return dim == 3 ? 64 / ft(n) : 128 / ft(n);
}
__host__ __device__ constexpr int z_rows(int n, int dim, int h) {
  return h && z_rows0(n, dim) >= 4 ? z_rows0(n, dim) / 2 : z_rows0(n, dim);
}
// first-axis projection pass: k-points per workgroup (same thread counts in
// FP64; FP32 takes twice as many for the same LDS and longer HBM runs)
__host__ __device__ constexpr int x_kpts0(int n, int dim, bool f32, int e) {
  // at most 1024 threads per workgroup
  return (dim == 3 ? 32 / ft(n) : 64 / ft(n)) * (f32 ? 2 : 1) * dim * dim * (n / e) <= 1024
             ? (dim == 3 ? 32 / ft(n) : 64 / ft(n)) * (f32 ? 2 : 1)
             : (dim == 3 ? 32 / ft(n) : 64 / ft(n));
}
// variant with half the k-points (h = 1): half the LDS per workgroup, so
// more workgroups per CU, for shorter HBM runs
__host__ __device__ constexpr int x_kpts(int n, int dim, bool f32, int e, int h) {
  return h && x_kpts0(n, dim, f32, e) >= 2 ? x_kpts0(n, dim, f32, e) / 2 : x_kpts0(n, dim, f32, e);
}
// LDS line stride: the line count rounded up to an odd number (lanes on
// consecutive lines then spread over the banks, and no padding is spent when
// the count is already odd)
__host__ __device__ constexpr int lstride(int L) { return L | 1; }
// strided passes: lines per workgroup (a run of 256 bytes)
template <typename T, int N, int E>
constexpr int s_lines() {
  return 256 / (int)sizeof(c_t<T>) * (N / E) <= 1024 ? 256 / (int)sizeof(c_t<T>) : 1024 / (N / E);
}

// ---- LDS per workgroup and occupancy ----------------------------------------
template <typename T, int N, int DIM, int H>
constexpr size_t lds_lines() {
  return (size_t)N * lstride(z_rows(N, DIM, H) / 2 * DIM * DIM) * sizeof(c_t<T>);
}
template <typename T, int N, int E, int DIM, int H>
constexpr size_t lds_c2r() {
  constexpr size_t lines = lds_lines<T, N, DIM, H>();
  constexpr size_t tile = (size_t)z_rows(N, DIM, H) * (N / 2 + 1) * DIM * DIM * sizeof(c_t<T>);
  constexpr size_t red = (size_t)z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E) * sizeof(double);
  return lines > tile ? (lines > red ? lines : red) : (tile > red ? tile : red);
}
template <typename T, int N, int E>
constexpr size_t lds_strided() {
  return (size_t)N * lstride(s_lines<T, N, E>()) * sizeof(c_t<T>);
}
template <typename T, int N, int E, int DIM, int H>
constexpr size_t lds_axis0() {
  return (size_t)N * lstride(x_kpts(N, DIM, sizeof(T) == 4, E, H) * DIM * DIM) * sizeof(c_t<T>);
}
// Twiddle table W_N^m, m < N, in LDS: a workgroup copies the table after its
// B bytes of lines, and the passes read it there (LDS reads, mostly
// broadcasts) instead of through the vector-memory path, where every
// butterfly's loads wait out the L1/L2 latency.  Only where the N extra
// values leave the number of workgroups the LDS holds per CU unchanged.
template <typename T, int N>
struct TwLds {
  static constexpr size_t tab = (size_t)N * sizeof(c_t<T>);
  static constexpr size_t off(size_t B) { return (B + 15) / 16 * 16; }
  static constexpr bool on(size_t B) { return 160 * 1024 / B == 160 * 1024 / (off(B) + tab); }
  static constexpr size_t total(size_t B) { return on(B) ? off(B) + tab : B; }
};
// the table the passes read: the LDS copy where TwLds::on(B) (every thread
// of the workgroup must call this), else tw itself
template <typename T, int N, int NT, size_t B>
__device__ __forceinline__ const c_t<T>* twiddles_for(const c_t<T>* __restrict__ tw,
                                                      unsigned char* smem) {
  if constexpr (TwLds<T, N>::on(B)) {
    c_t<T>* t = reinterpret_cast<c_t<T>*>(smem + TwLds<T, N>::off(B));
    for (int i = threadIdx.x; i < N; i += NT) t[i] = tw[i];
    __syncthreads();
    return t;
  } else {
    return tw;
  }
}
// Waves per SIMD the register allocation has to leave room for: two
// workgroups per CU wherever the 160 KiB of LDS hold two (one otherwise).
// Without the bound the compiler may take enough registers per lane that a
// second workgroup no longer fits, and a pass that alternates HBM phases
// with LDS phases then leaves HBM idle during the latter.
constexpr size_t kLdsPerCu = 160 * 1024;
__host__ __device__ constexpr int wpe(int threads, size_t lds) {
  return ((lds * 2 <= kLdsPerCu ? 2 : 1) * ((threads + 63) / 64) + 3) / 4;
}
#define MUGPU_WPE(threads, lds) __attribute__((amdgpu_waves_per_eu(wpe((threads), (lds)))))

// ---- strided c2c (middle axis, or first axis without projection) ----------
// data[(o * N + m) * inner + w]: workgroup (w-tile, o) transforms lines
// w0 .. w0 + L - 1 of plane o
template <typename T, int N, int E, int SG>
__device__ __forceinline__ void strided_body(c_t<T>* __restrict__ data, int64_t inner,
                                             const c_t<T>* __restrict__ tw, unsigned char* smem,
                                             int64_t bx, int64_t by, int64_t gx, int64_t gy) {
  constexpr int L = s_lines<T, N, E>(), LS = lstride(L);
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  const int l = threadIdx.x % L, p = threadIdx.x / L;
  const int64_t t = xcd_tile(by * gx + bx, gx * gy);
  const int64_t o = t / gx;
  const int64_t w = (t - o * gx) * L + l;
  const bool valid = w < inner;
  c_t<T>* base = data + o * N * inner + w;
  c_t<T> v[E];
  load_first<T, N, E>(v, p, [&](int m) { return valid ? base[m * inner] : czero<T>(); });
  Passes<T, N, E, SG, 0, 1>::run(v, p, tw, [&](int m) -> c_t<T>& { return s[m * LS + l]; });
  if (valid) store_last<T, N, E>(v, p, [&](int m, c_t<T> x) { base[m * inner] = x; });
}

template <typename T, int N, int E, int SG>
__global__ __launch_bounds__((s_lines<T, N, E>() * (N / E)))
MUGPU_WPE((s_lines<T, N, E>() * (N / E)), (lds_strided<T, N, E>())) void k_strided_fast(
    c_t<T>* __restrict__ data, int64_t inner, const c_t<T>* __restrict__ tw, const int* active) {
  if (active && !active[0]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const c_t<T>* t = twiddles_for<T, N, s_lines<T, N, E>() * (N / E), lds_strided<T, N, E>()>(tw, smem);
  strided_body<T, N, E, SG>(data, inner, t, smem, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y);
}

// batched (FP32, one RVE per blockIdx.z)
template <int N, int E, int SG>
__global__ __launch_bounds__((s_lines<float, N, E>() * (N / E)))
MUGPU_WPE((s_lines<float, N, E>() * (N / E)), (lds_strided<float, N, E>())) void k_strided_batch(
    const BatchSlot* __restrict__ slots, int64_t inner, const float2* __restrict__ tw) {
  const BatchSlot& sl = slots[blockIdx.z];
  if (!sl.ctl[CTL_ACTIVE]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const float2* t = twiddles_for<float, N, s_lines<float, N, E>() * (N / E), lds_strided<float, N, E>()>(tw, smem);
  strided_body<float, N, E, SG>(sl.spec, inner, t, smem, blockIdx.x, blockIdx.y, gridDim.x,
                             gridDim.y);
}

// ---- first axis: forward c2c, projection, backward c2c --------------------
template <typename T>
struct XArgs {
  c_t<T>* data;
  int64_t plane;  // k-points per first-axis index
  const c_t<T>* tw;
  int kind;
  int allow_even;
  int res[3];
  int hlast;
  double L[3];
  T norm;
  const int* active;  // FftPlan::active (single-cell kernels only)
};

template <typename T, int N, int E, int DIM, int H>
__device__ __forceinline__ void axis0_body(const XArgs<T>& a, unsigned char* smem, int64_t bx,
                                           int64_t gx) {
  constexpr int NC = DIM * DIM, KP = x_kpts(N, DIM, sizeof(T) == 4, E, H), L = KP * NC, LS = lstride(L),
                NT = L * (N / E);
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  const int l = threadIdx.x % L, p = threadIdx.x / L;
  const int64_t q0 = xcd_tile(bx, gx) * KP;
  const int64_t inner = a.plane * NC;
  const int64_t w = q0 * NC + l;
  const bool valid = w < inner;
  c_t<T>* base = a.data + w;
  auto at = [&](int m) -> c_t<T>& { return s[m * LS + l]; };
  c_t<T> v[E];
  load_first<T, N, E>(v, p, [&](int m) { return valid ? base[m * inner] : czero<T>(); });
  Passes<T, N, E, -1, 0, 1>::run(v, p, a.tw, at);
  store_last<T, N, E>(v, p, [&](int m, c_t<T> x) { at(m) = x; });
  __syncthreads();
  // projection per wave vector (k0, q0 + kp); NT is a multiple of KP, so a
  // thread keeps one kp and its (k1, k2) for every k0 it visits (a plane
  // holds fewer than 2^31 k-points)
  const int kp = threadIdx.x % KP;
  const int64_t q = q0 + kp;
  int c[3] = {0, 0, 0};
  if (DIM == 3) {
    c[1] = (int)q / a.hlast;
    c[2] = (int)q - c[1] * a.hlast;
  } else {
    c[1] = (int)q;
  }
  for (int k0 = threadIdx.x / KP; k0 < N && q < a.plane; k0 += NT / KP) {
    c[0] = k0;
    T xi[DIM];
    c_t<T>* f = s + k0 * LS + kp * NC;
    if (!unit_wave_vector<T, DIM>(c, a.res, a.L, a.allow_even, xi)) {
#pragma unroll
      for (int r = 0; r < NC; ++r) f[r] = czero<T>();
      continue;
    }
    project_point<T, DIM>(f, 1, xi, a.kind, a.norm);
  }
  __syncthreads();
  load_first<T, N, E>(v, p, [&](int m) { return at(m); });
  __syncthreads();
  Passes<T, N, E, +1, 0, 1>::run(v, p, a.tw, at);
  if (valid) store_last<T, N, E>(v, p, [&](int m, c_t<T> x) { base[m * inner] = x; });
}

template <typename T, int N, int E, int DIM, int H>
__device__ __forceinline__ void axis0_entry(const XArgs<T>& a, unsigned char* smem) {
  if (a.active && !a.active[0]) return;
  XArgs<T> b = a;
  b.tw = twiddles_for<T, N, x_kpts(N, DIM, sizeof(T) == 4, E, H) * DIM * DIM * (N / E),
                      lds_axis0<T, N, E, DIM, H>()>(a.tw, smem);
  axis0_body<T, N, E, DIM, H>(b, smem, blockIdx.x, gridDim.x);
}

template <typename T, int N, int E, int DIM, int H>
// (no register bound: the projection stage needs its registers, and spilling
// them costs more than the second workgroup per CU gains)
__global__ __launch_bounds__(x_kpts(N, DIM, sizeof(T) == 4, E, H) * DIM * DIM * (N / E)) void k_axis0_fast(XArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  axis0_entry<T, N, E, DIM, H>(a, smem);
}

// The same pass with the register allocation told to leave room for
// axis0_hint<T>() waves per SIMD (variant bit 2).  Where the LDS already
// caps the workgroups per CU, as at 512, the bound changes the register
// allocation and the instruction schedule, not the occupancy; there (E = 16)
// it shortens the FP64 pass by 17 % and the FP32 one by 3 %, with the same
// results bit for bit.  The tuned table chooses it per length.
template <typename T>
__host__ __device__ constexpr int axis0_hint() {
  return sizeof(T) == 8 ? 4 : 6;
}
template <typename T, int N, int E, int DIM, int H>
__global__ __launch_bounds__(x_kpts(N, DIM, sizeof(T) == 4, E, H) * DIM * DIM * (N / E))
__attribute__((amdgpu_waves_per_eu(axis0_hint<T>()))) void k_axis0_bounded(XArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  axis0_entry<T, N, E, DIM, H>(a, smem);
}

template <int N, int E, int DIM, int H>
__global__ __launch_bounds__(x_kpts(N, DIM, true, E, H) * DIM * DIM * (N / E)) void k_axis0_batch(
    const BatchSlot* __restrict__ slots, XArgs<float> shape) {
  const BatchSlot& sl = slots[blockIdx.z];
  if (!sl.ctl[CTL_ACTIVE]) return;
  XArgs<float> a = shape;
  a.data = sl.spec;
  a.kind = sl.kind;
  a.allow_even = sl.allow_even;
  for (int d = 0; d < 3; ++d) a.L[d] = sl.L[d];
  a.norm = (float)sl.norm;
  extern __shared__ __align__(16) unsigned char smem[];
  a.tw = twiddles_for<float, N, x_kpts(N, DIM, true, E, H) * DIM * DIM * (N / E),
                      lds_axis0<float, N, E, DIM, H>()>(shape.tw, smem);
  axis0_body<float, N, E, DIM, H>(a, smem, blockIdx.x, gridDim.x);
}

// ---- last axis: input stage, r2c ---------------------------------------------
// Rows 2 rp and 2 rp + 1 of component c share complex line rp * NC + c (row
// 2 rp in the real part), separated after the transform by Hermitian symmetry:
//   X_a[k] = (Z[k] + conj Z[n-k]) / 2,  X_b[k] = (Z[k] - conj Z[n-k]) / 2i
template <typename T>
struct LArgs {
  const T* in;
  const T* rvec;
  T* pout;
  const double* scal;
  const T* tan;
  int tan_kind;
  c_t<T>* spec;
  T* out;
  const T* dotv;
  double* partials;
  int64_t rows;
  const c_t<T>* tw;
  const int* active;  // FftPlan::active (single-cell kernels only)
  T* x;               // OpSpec::xvec (IN_TAN_PUPDATE_X kernels only)
};

// ISO: the tangent is known to be isotropic (TAN_ISO), which the compiler
// folds in; otherwise a.tan_kind selects it at run time
template <typename T, int N, int E, int DIM, int IN, bool ISO, int H>
__device__ __forceinline__ void r2c_body(const LArgs<T>& a, unsigned char* smem, int64_t bx) {
  constexpr int NC = DIM * DIM, RB = z_rows(N, DIM, H), L = RB / 2 * NC, LS = lstride(L);
  constexpr int NT = L * (N / E), NH = N / 2 + 1;
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  T* sr = reinterpret_cast<T*>(smem);
  const int64_t row0 = bx * RB;
  if constexpr (IN == IN_PLAIN) {
    const int64_t base = row0 * N * NC, lim = a.rows * N * NC;
    for (int e = threadIdx.x; e < RB * N * NC; e += NT) {
      const int r = e / (N * NC), rem = e - r * (N * NC);
      const int m = rem / NC, c = rem - m * NC;
      sr[2 * (m * LS + (r >> 1) * NC + c) + (r & 1)] = base + e < lim ? a.in[base + e] : T(0);
    }
  } else {
    // p (updated on the way in: p = -r + beta p) is staged in LDS by an
    // element-wise sweep, consecutive lanes on consecutive values, so that
    // the AoS fields are read and written in full cache lines; clamped
    // unconditional loads, kBatch per thread in flight.
    constexpr bool PU = IN == IN_TAN_PUPDATE || IN == IN_TAN_PUPDATE_X;
    // the previous iteration's x update, x += alpha p_old, when fused here
    constexpr bool XU = IN == IN_TAN_PUPDATE_X;
    const T beta = PU ? (T)a.scal[SC_BETA] : T(0);
    const T al = XU ? (T)a.scal[SC_ALPHA] : T(0);
    {
      constexpr int TOT = RB * N * NC, kBatch = 8;
      const int64_t base = row0 * N * NC, lim = a.rows * N * NC;
      for (int e0 = 0; e0 < TOT; e0 += kBatch * NT) {
        T pv[kBatch], rv[kBatch], xv[kBatch];
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
          const int e = e0 + i * NT + threadIdx.x;
          const int64_t g = base + e < lim ? base + e : lim - 1;
          pv[i] = a.in[g];
          if (PU) rv[i] = a.rvec[g];
          if (XU) xv[i] = a.x[g];
        }
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
          const int e = e0 + i * NT + threadIdx.x;
          if (e < TOT) {
            const bool valid = base + e < lim;
            T x = T(0);
            if (valid) {
              if (PU) {
                if (XU) a.x[base + e] = xv[i] + al * pv[i];
                x = -rv[i] + beta * pv[i];
                a.pout[base + e] = x;
              } else {
                x = pv[i];
              }
            }
            const int r = e / (N * NC), rem = e - r * (N * NC);
            const int m = rem / NC, c = rem - m * NC;
            sr[2 * (m * LS + (r >> 1) * NC + c) + (r & 1)] = x;
          }
        }
      }
    }
    __syncthreads();
    // K:p per voxel, in place: each voxel's components are read and written
    // by one thread.  One loop per tangent kind, the kind a compile-time
    // constant in each: a run-time switch inside the loop keeps the other
    // kinds' values live and leaves too few registers for the tangent loads,
    // which then go out one at a time.
    auto kp_loop = [&](auto kind_c) {
      constexpr int KIND = decltype(kind_c)::value;
      for (int e = threadIdx.x; e < RB * N; e += NT) {
        const int r = e / N, m = e - (e / N) * N;
        const int64_t row = row0 + r;
        if (row < a.rows) {
          T pv[NC], dp[NC];
#pragma unroll
          for (int c = 0; c < NC; ++c) pv[c] = sr[2 * (m * LS + (r >> 1) * NC + c) + (r & 1)];
          tangent_product<T, DIM>(KIND, a.tan, row * N + m, pv, dp);
#pragma unroll
          for (int c = 0; c < NC; ++c) sr[2 * (m * LS + (r >> 1) * NC + c) + (r & 1)] = dp[c];
        }
      }
    };
    if constexpr (ISO) {
      kp_loop(std::integral_constant<int, TAN_ISO>{});
    } else {
      switch (a.tan_kind) {
        case TAN_SYM: kp_loop(std::integral_constant<int, TAN_SYM>{}); break;
        case TAN_HEP: kp_loop(std::integral_constant<int, TAN_HEP>{}); break;
        case TAN_ISO: kp_loop(std::integral_constant<int, TAN_ISO>{}); break;
        default: kp_loop(std::integral_constant<int, TAN_FULL>{}); break;
      }
    }
  }
  __syncthreads();
  const int l = threadIdx.x % L, p = threadIdx.x / L;
  auto at = [&](int m) -> c_t<T>& { return s[m * LS + l]; };
  c_t<T> v[E];
  load_first<T, N, E>(v, p, [&](int m) { return at(m); });
  __syncthreads();
  Passes<T, N, E, -1, 0, 1>::run(v, p, a.tw, at);
  store_last<T, N, E>(v, p, [&](int m, c_t<T> x) { at(m) = x; });
  __syncthreads();
  const int64_t obase = row0 * NH * NC, olim = a.rows * NH * NC;
  for (int e = threadIdx.x; e < RB * NH * NC; e += NT) {
    if (obase + e >= olim) break;
    const int r = e / (NH * NC), rem = e - r * (NH * NC);
    const int k = rem / NC, c = rem - k * NC;
    const int ln = (r >> 1) * NC + c;
    const c_t<T> z = s[k * LS + ln];
    const c_t<T> zc = s[(k == 0 ? 0 : N - k) * LS + ln];
    c_t<T> X;
    if (!(r & 1)) {
      X.x = T(0.5) * (z.x + zc.x);
      X.y = T(0.5) * (z.y - zc.y);
    } else {
      X.x = T(0.5) * (z.y + zc.y);
      X.y = T(-0.5) * (z.x - zc.x);
    }
    a.spec[obase + e] = X;
  }
}

template <typename T, int N, int E, int DIM, int IN, bool ISO, int H>
__global__ __launch_bounds__(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E))
MUGPU_WPE(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), (lds_lines<T, N, DIM, H>())) void k_r2c_fast(LArgs<T> a) {
  if (a.active && !a.active[0]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  LArgs<T> b = a;
  b.tw = twiddles_for<T, N, z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), lds_lines<T, N, DIM, H>()>(a.tw, smem);
  r2c_body<T, N, E, DIM, IN, ISO, H>(b, smem, blockIdx.x);
}

// batched: the first iteration of a solve takes p as it is, later ones
// update it (p = -r + beta p) on the way in and apply the previous
// iteration's x update (x += alpha p_old), as the single-cell solve does;
// the served material is isotropic (TAN_ISO tangents)
template <int N, int E, int DIM, int H>
__global__ __launch_bounds__(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E))
MUGPU_WPE(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), (lds_lines<float, N, DIM, H>())) void k_r2c_batch(
    const BatchSlot* __restrict__ slots, int64_t rows, const float2* __restrict__ tw) {
  const BatchSlot& sl = slots[blockIdx.z];
  if (!sl.ctl[CTL_ACTIVE]) return;
  LArgs<float> a{};
  a.in = sl.p;
  a.rvec = sl.r;
  a.pout = sl.p;
  a.scal = sl.scal;
  a.tan = sl.tan;
  a.tan_kind = sl.tan_kind;
  a.spec = sl.spec;
  a.rows = rows;
  a.x = sl.x;
  extern __shared__ __align__(16) unsigned char smem[];
  a.tw = twiddles_for<float, N, z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), lds_lines<float, N, DIM, H>()>(tw, smem);
  if (sl.ctl[CTL_ITER] == 0) r2c_body<float, N, E, DIM, IN_TAN, true, H>(a, smem, blockIdx.x);
  else r2c_body<float, N, E, DIM, IN_TAN_PUPDATE_X, true, H>(a, smem, blockIdx.x);
}

// ---- last axis: c2r, output stage ---------------------------------------------
// Half lines A (row 2 rp) and B (row 2 rp + 1) are completed by Hermitian
// symmetry (imaginary parts of k = 0 and k = n/2 discarded) and combined into
// Z = A + i B, whose backward transform holds row 2 rp in its real and row
// 2 rp + 1 in its imaginary part.
template <typename T, int N, int E, int DIM, int OUT, int H>
__device__ __forceinline__ void c2r_body(const LArgs<T>& a, unsigned char* smem, int64_t bx) {
  constexpr int NC = DIM * DIM, RB = z_rows(N, DIM, H), L = RB / 2 * NC, LS = lstride(L);
  constexpr int NT = L * (N / E), NH = N / 2 + 1;
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  const int64_t row0 = bx * RB;
  {
    // loads issued in batches of kBatch before their LDS stores, so that each
    // thread keeps several HBM requests in flight
    constexpr int TOT = RB * NH * NC, kBatch = 8;
    const int64_t base = row0 * NH * NC, lim = a.rows * NH * NC;
    for (int e0 = 0; e0 < TOT; e0 += kBatch * NT) {
      c_t<T> t[kBatch];
      // clamped, unconditional loads (a predicated load per element would be
      // compiled as a branch that waits for it); out-of-range values are
      // replaced afterwards
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int e = e0 + i * NT + threadIdx.x;
        const int64_t g = base + e < lim ? base + e : lim - 1;
        t[i] = a.spec[g];
      }
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int e = e0 + i * NT + threadIdx.x;
        if (!(e < TOT && base + e < lim)) t[i] = czero<T>();
      }
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int e = e0 + i * NT + threadIdx.x;
        if (e < TOT) s[e] = t[i];
      }
    }
  }
  __syncthreads();
  const int l = threadIdx.x % L, p = threadIdx.x / L;
  const int rp = l / NC, cc = l - rp * NC;
  c_t<T> v[E];
  load_first<T, N, E>(v, p, [&](int i) {
    const int kk = i <= N / 2 ? i : N - i;
    c_t<T> A = s[((2 * rp) * NH + kk) * NC + cc];
    c_t<T> B = s[((2 * rp + 1) * NH + kk) * NC + cc];
    if (kk == 0 || 2 * kk == N) {
      A.y = T(0);
      B.y = T(0);
    }
    if (i > N / 2) {
      A.y = -A.y;
      B.y = -B.y;
    }
    c_t<T> z;
    z.x = A.x - B.y;
    z.y = A.y + B.x;
    return z;
  });
  __syncthreads();
  auto at = [&](int m) -> c_t<T>& { return s[m * LS + l]; };
  Passes<T, N, E, +1, 0, 1>::run(v, p, a.tw, at);
  store_last<T, N, E>(v, p, [&](int m, c_t<T> x) { at(m) = x; });
  __syncthreads();
  const T* sr = reinterpret_cast<const T*>(smem);
  const int64_t base = row0 * N * NC, lim = a.rows * N * NC;
  double acc = 0.0;
  {
    constexpr int TOT = RB * N * NC, kBatch = 8;
    for (int e0 = 0; e0 < TOT; e0 += kBatch * NT) {
      T dv[kBatch];
      if (OUT == OUT_DOT) {
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
          const int e = e0 + i * NT + threadIdx.x;
          const int64_t g = base + e < lim ? base + e : lim - 1;
          dv[i] = a.dotv[g];
        }
      }
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int e = e0 + i * NT + threadIdx.x;
        if (e < TOT && base + e < lim) {
          const int r = e / (N * NC), rem = e - r * (N * NC);
          const int m = rem / NC, c = rem - m * NC;
          const T val = sr[2 * (m * LS + (r >> 1) * NC + c) + (r & 1)];
          a.out[base + e] = val;
          if (OUT == OUT_DOT) acc += (double)val * (double)dv[i];
        }
      }
    }
  }
  if (OUT == OUT_DOT) {
    __syncthreads();
    const double t = block_sum_any<NT>(acc, reinterpret_cast<double*>(smem));
    if (threadIdx.x == 0) a.partials[bx] = t;
  }
}

template <typename T, int N, int E, int DIM, int OUT, int H>
__global__ __launch_bounds__(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E))
MUGPU_WPE(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), (lds_c2r<T, N, E, DIM, H>())) void k_c2r_fast(LArgs<T> a) {
  if (a.active && !a.active[0]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  LArgs<T> b = a;
  b.tw = twiddles_for<T, N, z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), lds_c2r<T, N, E, DIM, H>()>(a.tw, smem);
  c2r_body<T, N, E, DIM, OUT, H>(b, smem, blockIdx.x);
}

template <int N, int E, int DIM, int H>
__global__ __launch_bounds__(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E))
MUGPU_WPE(z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), (lds_c2r<float, N, E, DIM, H>())) void k_c2r_batch(
    const BatchSlot* __restrict__ slots, int64_t rows, const float2* __restrict__ tw) {
  const BatchSlot& sl = slots[blockIdx.z];
  if (!sl.ctl[CTL_ACTIVE]) return;
  LArgs<float> a{};
  a.spec = sl.spec;
  a.out = sl.Ap;
  a.dotv = sl.p;
  a.partials = sl.zpart;
  a.rows = rows;
  extern __shared__ __align__(16) unsigned char smem[];
  a.tw = twiddles_for<float, N, z_rows(N, DIM, H) / 2 * DIM * DIM * (N / E), lds_c2r<float, N, E, DIM, H>()>(tw, smem);
  c2r_body<float, N, E, DIM, OUT_DOT, H>(a, smem, blockIdx.x);
}

// ---- launchers ----------------------------------------------------------------
template <typename T>
const c_t<T>* tw_of(const FftPlan& p, int d) {
  return reinterpret_cast<const c_t<T>*>(sizeof(T) == 8 ? p.tw64[d] : p.tw32[d]);
}

// dispatch on the axis length and the variant (0: E = 16, or 12 for the
// lengths with a factor 3; 1: E = 8, powers of two only)
#define MUGPU_FAST_SWITCH(n, v, CALL)                  \
  switch (n) {                                         \
    case 16: return (v) ? CALL(16, 8) : CALL(16, 16);  \
    case 32: return (v) ? CALL(32, 8) : CALL(32, 16);  \
    case 48: return CALL(48, 12);                      \
    case 64: return (v) ? CALL(64, 8) : CALL(64, 16);  \
    case 96: return CALL(96, 12);                      \
    case 128: return (v) ? CALL(128, 8) : CALL(128, 16); \
    case 256: return (v) ? CALL(256, 8) : CALL(256, 16); \
    case 512: return (v) ? CALL(512, 8) : CALL(512, 16); \
    default: break;                                    \
  }

template <typename T>
constexpr int pi_of() {
  return sizeof(T) == 8 ? 0 : 1;
}

}  // namespace
}  // namespace mugpu
