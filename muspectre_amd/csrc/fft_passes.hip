// FFT passes of the operator G * (K : p) over AoS fields with nc interleaved
// components (the reference's FFTW plan: rank d, howmany = nc, stride nc,
// src/fft/fftw_engine.cc:59-110), fused with the stages around them:
//
//   last axis (contiguous):  r2c, fused with the tangent product K:p and
//                            the CG direction update p = -r + beta p
//   middle axis (3D only):   c2c forward / backward, strided
//   first axis:              c2c forward, Green-operator projection per
//                            wave vector, c2c backward, in one kernel
//   last axis:               c2r, fused with the p . Ap partial dot product
//
// Fields stay in the reference's AoS layout (nc components per pixel), so a
// tile of consecutive rows (last-axis passes) or of consecutive k-points
// (first-axis pass) is one contiguous block of HBM: every load and store is
// coalesced and every component of a voxel or wave vector lands in the same
// workgroup.  Spectra use the reference's Hermitian layout
// (N0, .., N_{d-1}/2+1) x nc complex.
#include <cmath>
#include <cstdio>
#include <type_traits>

#include "mugpu_internal.hpp"
#include "op_stages.hpp"

namespace mugpu {

template <typename T>
struct ZArgs {
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
  int rb;
  int ls;
  AxisPlan ap;
  const c_t<T>* tw;
  const int* active;  // FftPlan::active
};

template <int NT>
__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// ---- last axis, r2c -----------------------------------------------------------
// Real lines are paired (line 2q in the real part, 2q+1 in the imaginary
// part of one complex transform) and separated by Hermitian symmetry:
//   X_a[k] = (Z[k] + conj Z[n-k]) / 2,  X_b[k] = (Z[k] - conj Z[n-k]) / 2i
template <typename T, int NT, int DIM, int IN>
__global__ __launch_bounds__(NT) void k_last_r2c(ZArgs<T> a) {
  if (a.active && !a.active[0]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  T* sr = reinterpret_cast<T*>(smem);
  constexpr int NC = DIM * DIM;
  const int n = a.ap.n;
  const int rb = a.rb;
  const int ls = a.ls;
  const int64_t row0 = (int64_t)blockIdx.x * rb;
  const int nl2 = (rb * NC + 1) / 2;
  if (IN == IN_PLAIN) {
    const int tot = rb * n * NC;
    const int64_t base = row0 * n * NC;
    const int64_t lim = a.rows * n * NC;
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int r = e / (n * NC);
      const int rem = e - r * n * NC;
      const int m = rem / NC;
      const int c = rem - m * NC;
      const T v = (base + e < lim) ? a.in[base + e] : T(0);
      const int L = r * NC + c;
      sr[2 * ((L >> 1) * ls + m) + (L & 1)] = v;
    }
  } else {
    const int tot = rb * n;
    const T beta = (IN == IN_TAN_PUPDATE) ? (T)a.scal[SC_BETA] : T(0);
    // one loop per tangent kind (the kind a compile-time constant in each,
    // so the other kinds hold no registers inside the loop), and every load
    // of a voxel ahead of its first store of p (p is updated in place)
    auto voxels = [&](auto kind_c) {
      constexpr int KIND = decltype(kind_c)::value;
      for (int e = threadIdx.x; e < tot; e += NT) {
        const int r = e / n;
        const int m = e - r * n;
        const int64_t row = row0 + r;
        T dp[NC];
        if (row < a.rows) {
          const int64_t vox = row * n + m;
          T p[NC];
          if (IN == IN_TAN_PUPDATE) {
            T rv[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
              rv[c] = a.rvec[vox * NC + c];
              p[c] = a.in[vox * NC + c];
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) {
              p[c] = -rv[c] + beta * p[c];
              a.pout[vox * NC + c] = p[c];
            }
          } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) p[c] = a.in[vox * NC + c];
          }
          tangent_product<T, DIM>(KIND, a.tan, vox, p, dp);
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c) dp[c] = T(0);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int L = r * NC + c;
          sr[2 * ((L >> 1) * ls + m) + (L & 1)] = dp[c];
        }
      }
    };
    switch (a.tan_kind) {
      case TAN_SYM: voxels(std::integral_constant<int, TAN_SYM>{}); break;
      case TAN_HEP: voxels(std::integral_constant<int, TAN_HEP>{}); break;
      case TAN_ISO: voxels(std::integral_constant<int, TAN_ISO>{}); break;
      default: voxels(std::integral_constant<int, TAN_FULL>{}); break;
    }
  }
  if ((rb * NC) & 1) {
    const int L = rb * NC;
    for (int m = threadIdx.x; m < n; m += NT) sr[2 * ((L >> 1) * ls + m) + 1] = T(0);
  }
  __syncthreads();
  lds_fft<T, NT>(s, nl2, ls, a.ap, a.tw, -1);
  const int nh = n / 2 + 1;
  const int tot = rb * nh * NC;
  const int64_t obase = row0 * nh * NC;
  const int64_t olim = a.rows * nh * NC;
  for (int e = threadIdx.x; e < tot; e += NT) {
    if (obase + e >= olim) break;
    const int r = e / (nh * NC);
    const int rem = e - r * nh * NC;
    const int k = rem / NC;
    const int c = rem - k * NC;
    const int L = r * NC + c;
    const c_t<T>* line = s + (L >> 1) * ls;
    const c_t<T> z = line[k];
    const c_t<T> zc = line[k == 0 ? 0 : n - k];
    c_t<T> X;
    if (!(L & 1)) {
      X.x = T(0.5) * (z.x + zc.x);
      X.y = T(0.5) * (z.y - zc.y);
    } else {
      X.x = T(0.5) * (z.y + zc.y);
      X.y = T(-0.5) * (z.x - zc.x);
    }
    a.spec[obase + e] = X;
  }
}

// ---- last axis, c2r -------------------------------------------------------------
// Two Hermitian half-lines A, B are completed and combined into Z = A + iB
// (imaginary parts of k = 0 and, for even n, k = n/2 discarded); the
// backward transform of Z holds line a in its real and line b in its
// imaginary part.
template <typename T, int NT, int DIM, int OUT>
__global__ __launch_bounds__(NT) void k_last_c2r(ZArgs<T> a) {
  if (a.active && !a.active[0]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NT];
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  constexpr int NC = DIM * DIM;
  const int n = a.ap.n;
  const int nh = n / 2 + 1;
  const int rb = a.rb;
  const int ls = a.ls;
  const int64_t row0 = (int64_t)blockIdx.x * rb;
  const int nl2 = (rb * NC + 1) / 2;
  {
    const int tot = rb * nh * NC;
    const int64_t base = row0 * nh * NC;
    const int64_t lim = a.rows * nh * NC;
    for (int e = threadIdx.x; e < tot; e += NT) {
      c_t<T> v;
      if (base + e < lim) v = a.spec[base + e];
      else { v.x = T(0); v.y = T(0); }
      s[e] = v;
    }
  }
  __syncthreads();
  c_t<T> z[kEPT];
  const int nz = nl2 * n;
  const int nlines = rb * NC;
#pragma unroll
  for (int i = 0; i < kEPT; ++i) {
    const int idx = threadIdx.x + i * NT;
    if (idx < nz) {
      const int line = idx / n;
      const int k = idx - line * n;
      const int La = 2 * line, Lb = 2 * line + 1;
      const int kk = k < nh ? k : n - k;
      c_t<T> A, B;
      A.x = A.y = B.x = B.y = T(0);
      if (La < nlines) A = s[(La / NC) * nh * NC + kk * NC + (La % NC)];
      if (Lb < nlines) B = s[(Lb / NC) * nh * NC + kk * NC + (Lb % NC)];
      if (kk == 0 || (2 * kk == n)) {
        A.y = T(0);
        B.y = T(0);
      }
      if (k >= nh) {
        A.y = -A.y;
        B.y = -B.y;
      }
      c_t<T> zz;
      zz.x = A.x - B.y;
      zz.y = A.y + B.x;
      z[i] = zz;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kEPT; ++i) {
    const int idx = threadIdx.x + i * NT;
    if (idx < nz) {
      const int line = idx / n;
      const int k = idx - line * n;
      s[line * ls + k] = z[i];
    }
  }
  __syncthreads();
  lds_fft<T, NT>(s, nl2, ls, a.ap, a.tw, +1);
  const T* sr = reinterpret_cast<const T*>(s);
  const int tot = rb * n * NC;
  const int64_t base = row0 * n * NC;
  const int64_t lim = a.rows * n * NC;
  double acc = 0.0;
  for (int e = threadIdx.x; e < tot; e += NT) {
    if (base + e >= lim) break;
    const int r = e / (n * NC);
    const int rem = e - r * n * NC;
    const int m = rem / NC;
    const int c = rem - m * NC;
    const int L = r * NC + c;
    const T v = sr[2 * ((L >> 1) * ls + m) + (L & 1)];
    a.out[base + e] = v;
    if (OUT == OUT_DOT) acc += (double)v * (double)a.dotv[base + e];
  }
  if (OUT == OUT_DOT) {
    __syncthreads();
    const double t = block_sum<NT>(acc, red);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = t;
  }
}

// ---- strided c2c along one axis -------------------------------------------------
// data[(o*n + m)*inner + w]: tile of W consecutive w for one o, all m
template <typename T>
struct SArgs {
  c_t<T>* data;
  int64_t inner;
  int W;
  int ls;
  int sign;
  AxisPlan ap;
  const c_t<T>* tw;
  const int* active;  // FftPlan::active
};

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_strided(SArgs<T> a) {
  if (a.active && !a.active[0]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  const int n = a.ap.n;
  const int64_t w0 = (int64_t)blockIdx.x * a.W;
  const int64_t rem = a.inner - w0;
  const int W = rem < a.W ? (int)rem : a.W;
  c_t<T>* base = a.data + (int64_t)blockIdx.y * n * a.inner + w0;
  const int tot = n * W;
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int m = e / W;
    const int wl = e - m * W;
    s[wl * a.ls + m] = base[(int64_t)m * a.inner + wl];
  }
  __syncthreads();
  lds_fft<T, NT>(s, W, a.ls, a.ap, a.tw, a.sign);
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int m = e / W;
    const int wl = e - m * W;
    base[(int64_t)m * a.inner + wl] = s[wl * a.ls + m];
  }
}

// ---- first axis: forward c2c, projection, backward c2c --------------------------
template <typename T>
struct PArgs {
  c_t<T>* data;
  int64_t plane;  // k-points per first-axis index
  int Wk;
  int ls;
  AxisPlan ap;
  const c_t<T>* tw;
  int kind;
  int allow_even;
  int res[3];
  int hlast;
  double L[3];
  T norm;
  const int* active;  // FftPlan::active
};

template <typename T, int NT, int DIM>
__global__ __launch_bounds__(NT) void k_axis0_project(PArgs<T> a) {
  if (a.active && !a.active[0]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  c_t<T>* s = reinterpret_cast<c_t<T>*>(smem);
  constexpr int NC = DIM * DIM;
  const int n = a.ap.n;
  const int64_t q0 = (int64_t)blockIdx.x * a.Wk;
  const int64_t rem = a.plane - q0;
  const int Wk = rem < a.Wk ? (int)rem : a.Wk;
  const int W = Wk * NC;
  const int64_t inner = a.plane * NC;
  c_t<T>* base = a.data + q0 * NC;
  const int tot = n * W;
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int m = e / W;
    const int wl = e - m * W;
    s[wl * a.ls + m] = base[(int64_t)m * inner + wl];
  }
  __syncthreads();
  lds_fft<T, NT>(s, W, a.ls, a.ap, a.tw, -1);
  // projection per wave vector (src/fft/projection_small_strain.cc:61-85,
  // projection_default.cc:57-61, projection_finite_strain_fast.cc:65-85)
  const int tp = n * Wk;
  for (int e = threadIdx.x; e < tp; e += NT) {
    const int k0 = e / Wk;
    const int wk = e - k0 * Wk;
    const int64_t q = q0 + wk;
    int c[3];
    c[0] = k0;
    if (DIM == 3) {
      // a plane holds fewer than 2^31 k-points
      c[1] = (int)q / a.hlast;
      c[2] = (int)q - c[1] * a.hlast;
    } else {
      c[1] = (int)q;
    }
    T xi[DIM];
    c_t<T>* f = s + (wk * NC) * a.ls + k0;
    if (!unit_wave_vector<T, DIM>(c, a.res, a.L, a.allow_even, xi)) {
#pragma unroll
      for (int r = 0; r < NC; ++r) {
        f[r * a.ls].x = T(0);
        f[r * a.ls].y = T(0);
      }
      continue;
    }
    project_point<T, DIM>(f, a.ls, xi, a.kind, a.norm);
  }
  __syncthreads();
  lds_fft<T, NT>(s, W, a.ls, a.ap, a.tw, +1);
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int m = e / W;
    const int wl = e - m * W;
    base[(int64_t)m * inner + wl] = s[wl * a.ls + m];
  }
}

// ============================== host side ======================================
namespace {

constexpr int kNT = 256;

int factorize(int n, AxisPlan& ap) {
  ap.n = n;
  ap.nf = 0;
  int m = n;
  auto push = [&](int r) {
    if (ap.nf >= kMaxFactors) return false;
    ap.f[ap.nf++] = r;
    m /= r;
    return true;
  };
  while (m % 8 == 0)
    if (!push(8)) return MUGPU_EARG;
  while (m % 4 == 0)
    if (!push(4)) return MUGPU_EARG;
  while (m % 2 == 0)
    if (!push(2)) return MUGPU_EARG;
  for (int q = 3; m > 1; q += 2)
    while (m % q == 0)
      if (!push(q)) return MUGPU_EARG;
  return MUGPU_OK;
}

// exp(-2 pi i j / n) with the octant reduction of the angle
void twiddles(int n, std::vector<double>& out) {
  out.resize(2 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    const long a = 8L * j;
    const long oct = a / n;
    const double t = (double)(a - oct * n) / (double)n * (M_PI / 4.0);
    const double ct = std::cos(t), st = std::sin(t);
    const double cc = std::cos(M_PI / 4.0 - t), sc = std::sin(M_PI / 4.0 - t);
    double c, s;
    switch (oct) {
      case 0: c = ct; s = st; break;
      case 1: c = sc; s = cc; break;
      case 2: c = -st; s = ct; break;
      case 3: c = -cc; s = sc; break;
      case 4: c = -ct; s = -st; break;
      case 5: c = -sc; s = -cc; break;
      case 6: c = st; s = -ct; break;
      default: c = cc; s = -sc; break;
    }
    out[2 * j] = c;
    out[2 * j + 1] = -s;
  }
}

inline int odd_stride(int n) { return n | 1; }

template <typename T>
size_t csize() {
  return sizeof(c_t<T>);
}

}  // namespace

int plan_init(FftPlan& p, int dim, const int* res, int nc, int prec) {
  if (dim != 2 && dim != 3) {
    set_error("only 2D and 3D grids are supported");
    return MUGPU_EARG;
  }
  if (nc != dim * dim) {
    set_error("the HIP engine transforms nb_components = dim^2 = %d fields, got %d", dim * dim, nc);
    return MUGPU_EARG;
  }
  p.dim = dim;
  p.nc = nc;
  p.prec = prec;
  p.N = 1;
  for (int d = 0; d < dim; ++d) {
    if (res[d] < 1) {
      set_error("resolutions must be positive");
      return MUGPU_EARG;
    }
    p.res[d] = res[d];
    p.hres[d] = res[d];
    p.N *= res[d];
  }
  p.hres[dim - 1] = res[dim - 1] / 2 + 1;
  p.kpts = 1;
  for (int d = 0; d < dim; ++d) p.kpts *= p.hres[d];
  for (int d = 0; d < dim; ++d) {
    if (factorize(res[d], p.ax[d])) {
      set_error("too many prime factors in %d", res[d]);
      return MUGPU_EARG;
    }
    std::vector<double> tw;
    twiddles(res[d], tw);
    std::vector<float> tw32(tw.begin(), tw.end());
    MUGPU_TRY(dev_alloc(&p.tw64[d], tw.size() * sizeof(double), "the twiddle tables"));
    MUGPU_TRY(dev_alloc(&p.tw32[d], tw32.size() * sizeof(float), "the twiddle tables"));
    MUGPU_HIP(hipMemcpy(p.tw64[d], tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
    MUGPU_HIP(hipMemcpy(p.tw32[d], tw32.data(), tw32.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  // every pass keeps its lines in LDS (FP64 worst case: 16 B per complex,
  // line stride n|1) and at most kNT*kEPT complex elements in registers
  // headroom for the static reduction arrays (up to 1024 doubles)
  const size_t lds = (size_t)max_lds_per_block() - 8192;
  auto fits = [&](int64_t lines, int len, int nt) {
    return lines * len <= (int64_t)nt * kEPT &&
           (size_t)lines * odd_stride(len) * 16 <= lds;
  };
  for (int d = 0; d < dim; ++d) p.fastax[d] = fast_size(res[d]);
  // variants of the compile-time-length passes, from the tuned table
  for (int pr = 0; pr < 2; ++pr)
    for (int iso = 0; iso < 2; ++iso) {
      const int prec = pr == 0 ? MUGPU_FP64 : MUGPU_FP32;
      if (p.fastax[dim - 1]) {
        p.fvar[pr][FV_R2C][iso] = (unsigned char)fast_variant(FV_R2C, res[dim - 1], dim, prec, iso);
        p.fvar[pr][FV_C2R][iso] = (unsigned char)fast_variant(FV_C2R, res[dim - 1], dim, prec, 0);
      }
      if (p.fastax[0]) {
        p.fvar[pr][FV_AXIS0][iso] = (unsigned char)fast_variant(FV_AXIS0, res[0], dim, prec, 0);
        p.fvar[pr][FV_STRIDED0][iso] = (unsigned char)fast_variant(FV_STRIDED0, res[0], dim, prec, 0);
      }
      if (dim == 3 && p.fastax[1])
        p.fvar[pr][FV_STRIDED1][iso] = (unsigned char)fast_variant(FV_STRIDED1, res[1], dim, prec, 0);
    }
  const int n = res[dim - 1];
  const int64_t rows = p.N / n;
  if (p.fastax[dim - 1]) {
    for (int pr = 0; pr < 2; ++pr) {
      const int rb = fast_rows_per_block(n, dim, p.fvar[pr][FV_C2R][0]);
      p.nzb[pr] = (rows + rb - 1) / rb;
    }
    p.zrb = fast_rows_per_block(n, dim, 0);
  } else {
    if (!fits((nc + 1) / 2, n, kNT)) {
      set_error("last axis of %d points exceeds the supported line length", n);
      return MUGPU_EARG;
    }
    // rows per workgroup: as many as fit LDS (up to 64) while the grid still
    // has kMinBlocks workgroups; small grids get one row per workgroup
    constexpr int64_t kMinBlocks = 2048;
    int rb = 1;
    while (rb < rows && rb < 64 && rows / (rb + 1) >= kMinBlocks &&
           fits(((rb + 1) * nc + 1) / 2, n, kNT) &&
           fits((int64_t)(rb + 1) * (n / 2 + 1) * nc, 1, kNT * 64))
      ++rb;
    p.zrb = rb;
  }
  if (!p.fastax[dim - 1]) p.nzb[0] = p.nzb[1] = (rows + p.zrb - 1) / p.zrb;
  p.nzblocks = p.nzb[0] > p.nzb[1] ? p.nzb[0] : p.nzb[1];
  // strided middle axis: up to 16 lines (256 B rows in FP64)
  if (dim == 3 && !p.fastax[1]) {
    const int ny = res[1];
    int W = 16;
    while (W > 1 && !fits(W, ny, kNT)) W /= 2;
    if (!fits(W, ny, kNT)) {
      set_error("middle axis of %d points exceeds the supported line length", ny);
      return MUGPU_EARG;
    }
    p.yW = W;
  }
  const int nx = res[0];
  if (!p.fastax[0]) {
    int W = 16;
    while (W > 1 && !fits(W, nx, kNT)) W /= 2;
    if (!fits(W, nx, kNT)) {
      set_error("first axis of %d points exceeds the supported line length", nx);
      return MUGPU_EARG;
    }
    p.xW = W;
  }
  // fused first-axis pass: whole k-points (nc lines each)
  if (!p.fastax[0]) {
    int NT = 256;
    while (NT < 1024 && !fits(nc, nx, NT)) NT *= 2;
    if (!fits(nc, nx, NT)) {
      set_error("first axis of %d points exceeds the fused projection pass", nx);
      return MUGPU_EARG;
    }
    // two k-points per workgroup where they fit and the grid stays large
    int Wk = 2;
    if (p.kpts / nx < 2 * 2048) Wk = 1;
    while (Wk > 1 && !fits((int64_t)Wk * nc, nx, NT)) --Wk;
    p.xNT = NT;
    p.xWk = Wk;
  }
  // room for the c2r partials of every variant (the autotuner times them all
  // on one plan): the half-rows variant launches the most workgroups
  int64_t nzmax = p.nzblocks;
  if (p.fastax[dim - 1]) {
    const int rbmin = fast_rows_per_block(res[dim - 1], dim, 2);
    nzmax = (p.N / res[dim - 1] + rbmin - 1) / rbmin;
  }
  MUGPU_TRY(dev_alloc(&p.partials, sizeof(double) * (size_t)(nzmax > kRedBlocks ? nzmax : kRedBlocks),
                      "the transform's partial sums"));
  return MUGPU_OK;
}

void plan_free(FftPlan& p) {
  for (int d = 0; d < 3; ++d) {
    if (p.tw64[d]) (void)hipFree(p.tw64[d]);
    if (p.tw32[d]) (void)hipFree(p.tw32[d]);
    p.tw64[d] = p.tw32[d] = nullptr;
  }
  if (p.partials) (void)hipFree(p.partials);
  p.partials = nullptr;
}

namespace {

template <typename T>
const c_t<T>* tw_of(const FftPlan& p, int d) {
  return reinterpret_cast<const c_t<T>*>(sizeof(T) == 8 ? p.tw64[d] : p.tw32[d]);
}

template <typename T>
constexpr int prec_of() {
  return sizeof(T) == 8 ? MUGPU_FP64 : MUGPU_FP32;
}

template <typename T, int DIM>
int launch_r2c(FftPlan& p, const OpSpec& op, c_t<T>* spec) {
  if (p.fastax[DIM - 1]) {
    timing_begin(MUGPU_K_LAST_R2C);
    const int rc = fast_r2c(p, prec_of<T>(), op, spec);
    timing_end(MUGPU_K_LAST_R2C);
    return rc;
  }
  ZArgs<T> a{};
  a.in = reinterpret_cast<const T*>(op.in);
  a.rvec = reinterpret_cast<const T*>(op.rvec);
  a.pout = reinterpret_cast<T*>(op.pout);
  a.scal = op.scal;
  a.tan = reinterpret_cast<const T*>(op.tangent);
  a.tan_kind = op.tan_kind;
  a.spec = spec;
  const int n = p.res[DIM - 1];
  a.rows = p.N / n;
  a.active = p.active;
  a.rb = p.zrb;
  a.ls = odd_stride(n);
  a.ap = p.ax[DIM - 1];
  a.tw = tw_of<T>(p, DIM - 1);
  const int nl2 = (a.rb * DIM * DIM + 1) / 2;
  const size_t lds = (size_t)nl2 * a.ls * csize<T>();
  dim3 grid((unsigned)p.nzblocks);
  timing_begin(MUGPU_K_LAST_R2C);
  switch (op.in_mode) {
    case IN_PLAIN: hipLaunchKernelGGL((k_last_r2c<T, kNT, DIM, IN_PLAIN>), grid, dim3(kNT), lds, stream(), a); break;
    case IN_TAN: hipLaunchKernelGGL((k_last_r2c<T, kNT, DIM, IN_TAN>), grid, dim3(kNT), lds, stream(), a); break;
    default: hipLaunchKernelGGL((k_last_r2c<T, kNT, DIM, IN_TAN_PUPDATE>), grid, dim3(kNT), lds, stream(), a); break;
  }
  timing_end(MUGPU_K_LAST_R2C);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <typename T, int DIM>
int launch_c2r(FftPlan& p, const OpSpec& op, c_t<T>* spec) {
  if (p.fastax[DIM - 1]) {
    timing_begin(MUGPU_K_LAST_C2R);
    const int rc = fast_c2r(p, prec_of<T>(), op, spec);
    timing_end(MUGPU_K_LAST_C2R);
    return rc;
  }
  ZArgs<T> a{};
  a.spec = spec;
  a.out = reinterpret_cast<T*>(op.out);
  a.dotv = reinterpret_cast<const T*>(op.dotv);
  a.partials = reinterpret_cast<double*>(p.partials);
  const int n = p.res[DIM - 1];
  const int nh = n / 2 + 1;
  a.rows = p.N / n;
  a.active = p.active;
  a.rb = p.zrb;
  a.ls = odd_stride(n);
  a.ap = p.ax[DIM - 1];
  a.tw = tw_of<T>(p, DIM - 1);
  const int nc = DIM * DIM;
  const int nl2 = (a.rb * nc + 1) / 2;
  const size_t e1 = (size_t)nl2 * a.ls, e2 = (size_t)a.rb * nh * nc;
  const size_t lds = (e1 > e2 ? e1 : e2) * csize<T>();
  dim3 grid((unsigned)p.nzblocks);
  timing_begin(MUGPU_K_LAST_C2R);
  if (op.out_mode == OUT_DOT)
    hipLaunchKernelGGL((k_last_c2r<T, kNT, DIM, OUT_DOT>), grid, dim3(kNT), lds, stream(), a);
  else
    hipLaunchKernelGGL((k_last_c2r<T, kNT, DIM, OUT_PLAIN>), grid, dim3(kNT), lds, stream(), a);
  timing_end(MUGPU_K_LAST_C2R);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

// c2c along axis `ax` of the Hermitian grid
template <typename T>
int launch_strided(FftPlan& p, int ax, int W, c_t<T>* spec, int sign) {
  SArgs<T> a{};
  a.data = spec;
  int64_t inner = p.nc;
  for (int d = ax + 1; d < p.dim; ++d) inner *= p.hres[d];
  int64_t outer = 1;
  for (int d = 0; d < ax; ++d) outer *= p.hres[d];
  const int n = p.res[ax];
  if (n == 1) return MUGPU_OK;
  if (p.fastax[ax]) {
    const int kc = ax == 0 ? MUGPU_K_AXIS0 : MUGPU_K_MIDDLE;
    timing_begin(kc);
    const int rc = fast_strided(p, prec_of<T>(), ax, spec, sign);
    timing_end(kc);
    return rc;
  }
  a.inner = inner;
  a.active = p.active;
  a.W = W;
  a.ls = odd_stride(n);
  a.sign = sign;
  a.ap = p.ax[ax];
  a.tw = tw_of<T>(p, ax);
  const size_t lds = (size_t)W * a.ls * csize<T>();
  dim3 grid((unsigned)((inner + W - 1) / W), (unsigned)outer);
  const int kcls = ax == 0 ? MUGPU_K_AXIS0 : MUGPU_K_MIDDLE;
  timing_begin(kcls);
  hipLaunchKernelGGL((k_strided<T, kNT>), grid, dim3(kNT), lds, stream(), a);
  timing_end(kcls);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <typename T, int DIM>
int launch_project(FftPlan& p, const ProjParams& pj, c_t<T>* spec) {
  if (p.fastax[0]) {
    timing_begin(MUGPU_K_AXIS0);
    const int rc = fast_axis0_project(p, prec_of<T>(), pj, spec);
    timing_end(MUGPU_K_AXIS0);
    return rc;
  }
  PArgs<T> a{};
  a.data = spec;
  a.plane = p.kpts / p.res[0];
  a.Wk = p.xWk;
  const int n = p.res[0];
  a.ls = odd_stride(n);
  a.ap = p.ax[0];
  a.tw = tw_of<T>(p, 0);
  a.kind = pj.kind;
  a.allow_even = pj.allow_even;
  for (int d = 0; d < 3; ++d) {
    a.res[d] = p.res[d];
    a.L[d] = pj.L[d];
  }
  a.hlast = p.hres[DIM - 1];
  a.norm = (T)pj.norm;
  a.active = p.active;
  const size_t lds = (size_t)a.Wk * DIM * DIM * a.ls * csize<T>();
  dim3 grid((unsigned)((a.plane + a.Wk - 1) / a.Wk));
  timing_begin(MUGPU_K_AXIS0);
  switch (p.xNT) {
    case 256: hipLaunchKernelGGL((k_axis0_project<T, 256, DIM>), grid, dim3(256), lds, stream(), a); break;
    case 512: hipLaunchKernelGGL((k_axis0_project<T, 512, DIM>), grid, dim3(512), lds, stream(), a); break;
    default: hipLaunchKernelGGL((k_axis0_project<T, 1024, DIM>), grid, dim3(1024), lds, stream(), a); break;
  }
  timing_end(MUGPU_K_AXIS0);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <typename T, int DIM>
int operator_impl(FftPlan& p, const OpSpec& op, const ProjParams& pj, void* spectrum) {
  c_t<T>* spec = reinterpret_cast<c_t<T>*>(spectrum);
  // the fused x update is carried by the compiled last-axis r2c only
  if (op.xvec && !p.fastax[DIM - 1]) return MUGPU_EARG;
  MUGPU_TRY((launch_r2c<T, DIM>(p, op, spec)));
  if (DIM == 3) MUGPU_TRY(launch_strided<T>(p, 1, p.yW, spec, -1));
  if (pj.kind >= 0) {
    MUGPU_TRY((launch_project<T, DIM>(p, pj, spec)));
  } else {
    MUGPU_TRY(launch_strided<T>(p, 0, p.xW, spec, -1));
    MUGPU_TRY(launch_strided<T>(p, 0, p.xW, spec, +1));
  }
  if (DIM == 3) MUGPU_TRY(launch_strided<T>(p, 1, p.yW, spec, +1));
  return launch_c2r<T, DIM>(p, op, spec);
}

template <typename T, int DIM>
int r2c_impl(FftPlan& p, const void* in, void* spectrum) {
  c_t<T>* spec = reinterpret_cast<c_t<T>*>(spectrum);
  OpSpec op;
  op.in_mode = IN_PLAIN;
  op.in = in;
  MUGPU_TRY((launch_r2c<T, DIM>(p, op, spec)));
  if (DIM == 3) MUGPU_TRY(launch_strided<T>(p, 1, p.yW, spec, -1));
  return launch_strided<T>(p, 0, p.xW, spec, -1);
}

template <typename T, int DIM>
int c2r_impl(FftPlan& p, void* spectrum, void* out) {
  c_t<T>* spec = reinterpret_cast<c_t<T>*>(spectrum);
  MUGPU_TRY(launch_strided<T>(p, 0, p.xW, spec, +1));
  if (DIM == 3) MUGPU_TRY(launch_strided<T>(p, 1, p.yW, spec, +1));
  OpSpec op;
  op.out_mode = OUT_PLAIN;
  op.out = out;
  return launch_c2r<T, DIM>(p, op, spec);
}

}  // namespace

namespace {
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_tangent_product(int kind, const T* __restrict__ tan,
                                                         const T* __restrict__ in,
                                                         T* __restrict__ out, int64_t npix) {
  constexpr int NC = DIM * DIM;
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= npix) return;
  T p[NC], dp[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) p[c] = in[v * NC + c];
  tangent_product<T, DIM>(kind, tan, v, p, dp);
#pragma unroll
  for (int c = 0; c < NC; ++c) out[v * NC + c] = dp[c];
}
}  // namespace

// dP = K : dF per voxel (src/cell/cell_base.cc:182-188)
int run_tangent_product(int dim, int prec, int kind, const void* tan, const void* in, void* out,
                        int64_t npix) {
  dim3 grid((unsigned)((npix + 255) / 256));
  if (prec == MUGPU_FP64) {
    if (dim == 3)
      hipLaunchKernelGGL((k_tangent_product<double, 3>), grid, dim3(256), 0, stream(), kind,
                         (const double*)tan, (const double*)in, (double*)out, npix);
    else
      hipLaunchKernelGGL((k_tangent_product<double, 2>), grid, dim3(256), 0, stream(), kind,
                         (const double*)tan, (const double*)in, (double*)out, npix);
  } else {
    if (dim == 3)
      hipLaunchKernelGGL((k_tangent_product<float, 3>), grid, dim3(256), 0, stream(), kind,
                         (const float*)tan, (const float*)in, (float*)out, npix);
    else
      hipLaunchKernelGGL((k_tangent_product<float, 2>), grid, dim3(256), 0, stream(), kind,
                         (const float*)tan, (const float*)in, (float*)out, npix);
  }
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int run_operator(FftPlan& p, int prec, const OpSpec& op, const ProjParams& pj, void* spectrum) {
  if (prec == MUGPU_FP64)
    return p.dim == 3 ? operator_impl<double, 3>(p, op, pj, spectrum)
                      : operator_impl<double, 2>(p, op, pj, spectrum);
  return p.dim == 3 ? operator_impl<float, 3>(p, op, pj, spectrum)
                    : operator_impl<float, 2>(p, op, pj, spectrum);
}

int run_r2c(FftPlan& p, int prec, const void* in, void* spectrum) {
  if (prec == MUGPU_FP64)
    return p.dim == 3 ? r2c_impl<double, 3>(p, in, spectrum) : r2c_impl<double, 2>(p, in, spectrum);
  return p.dim == 3 ? r2c_impl<float, 3>(p, in, spectrum) : r2c_impl<float, 2>(p, in, spectrum);
}

int run_c2r(FftPlan& p, int prec, void* spectrum, void* out) {
  if (prec == MUGPU_FP64)
    return p.dim == 3 ? c2r_impl<double, 3>(p, spectrum, out) : c2r_impl<double, 2>(p, spectrum, out);
  return p.dim == 3 ? c2r_impl<float, 3>(p, spectrum, out) : c2r_impl<float, 2>(p, spectrum, out);
}

}  // namespace mugpu
