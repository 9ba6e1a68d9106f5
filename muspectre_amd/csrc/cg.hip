// Vector kernels of SolverCG (src/solver/solver_cg.cc:62-96) and of the
// Newton drivers (src/solver/solvers.cc:142-187).  Reductions are
// deterministic: a fixed grid writes per-workgroup partial sums, and one
// workgroup adds them in a fixed order, so a result does not depend on
// what else runs on the device.
#include "mugpu_internal.hpp"

namespace mugpu {

namespace {

constexpr int kRT = 256;

template <int NT>
__device__ double block_reduce(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

template <typename A, typename B>
__global__ __launch_bounds__(kRT) void k_dot(const A* __restrict__ a, const B* __restrict__ b,
                                             int64_t n, double* partials) {
  __shared__ double red[kRT];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride)
    acc += (double)a[i] * (double)b[i];
  const double t = block_reduce<kRT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void k_finalize(const double* __restrict__ partials,
                                                   int64_t count, double* scal, int mode,
                                                   int* ctl) {
  if (ctl && !ctl[CTL_ACTIVE]) return;
  __shared__ double red[1024];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += 1024) acc += partials[i];
  const double s = block_reduce<1024>(acc, red);
  if (threadIdx.x == 0) {
    switch (mode) {
      case FIN_ALPHA:
        scal[SC_PAP] = s;
        scal[SC_ALPHA] = scal[SC_RDR] / s;
        break;
      case FIN_BETA:
        scal[SC_NEWRDR] = s;
        scal[SC_BETA] = s / scal[SC_RDR];
        scal[SC_RDR] = s;
        if (ctl) {
          // SolverCG's loop test after this iteration; a NaN r.r stops the
          // loop without counting the iteration, as the host loop's break
          if (!(s == s)) {
            ctl[CTL_ACTIVE] = 0;
          } else {
            const int it = ctl[CTL_ITER] + 1;
            ctl[CTL_ITER] = it;
            if (!(it < ctl[CTL_MAXIT] && s > scal[SC_TOL2])) ctl[CTL_ACTIVE] = 0;
          }
        }
        break;
      case FIN_STORE_TMP0: scal[SC_TMP0] = s; break;
      case FIN_STORE_TMP1: scal[SC_TMP1] = s; break;
      default: scal[SC_RDR] = s; break;
    }
  }
}

// x += alpha p ; r += alpha Ap ; partial r.r
template <typename T>
__global__ __launch_bounds__(kRT) void k_cg_update(T* __restrict__ x, T* __restrict__ r,
                                                   const T* __restrict__ p,
                                                   const T* __restrict__ Ap,
                                                   const double* __restrict__ scal, int64_t n,
                                                   double* partials, const int* ctl) {
  if (ctl && !ctl[CTL_ACTIVE]) return;
  __shared__ double red[kRT];
  const double alpha = scal[SC_ALPHA];
  const T al = (T)alpha;
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    if (x) x[i] += al * p[i];
    const T rv = r[i] + al * Ap[i];
    r[i] = rv;
    acc += (double)rv * (double)rv;
  }
  const double t = block_reduce<kRT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// x += alpha p, alpha = scal[SC_ALPHA]
template <typename T>
__global__ __launch_bounds__(kRT) void k_axpy_alpha(T* __restrict__ x, const T* __restrict__ p,
                                                    const double* __restrict__ scal, int64_t n,
                                                    int slot) {
  const T al = (T)scal[slot];
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) x[i] += al * p[i];
}

// r = Ax - b ; p = -r (Ax == nullptr: x = 0, where A x is zero exactly)
template <typename T>
__global__ __launch_bounds__(kRT) void k_cg_init(const T* __restrict__ Ax, const T* __restrict__ b,
                                                 T* __restrict__ r, T* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const T rv = (Ax ? Ax[i] : T(0)) - b[i];
    r[i] = rv;
    p[i] = -rv;
  }
}

template <typename T>
__global__ __launch_bounds__(kRT) void k_axpby(double a, const T* __restrict__ x, double b,
                                               T* __restrict__ y, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kRT;
  const T ta = (T)a, tb = (T)b;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride)
    y[i] = ta * x[i] + tb * y[i];
}

template <typename S, typename D>
__global__ __launch_bounds__(kRT) void k_convert(const S* __restrict__ s, D* __restrict__ d,
                                                 int64_t n, double scale) {
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride)
    d[i] = (D)((double)s[i] * scale);
}

// start of an inner FP32 solve on A d = r64 / |r64| (s = 1/|r64|):
// r32 = -s r64, p32 = -r32, x32 = 0 and the partials of r32.r32, reading r64
// once (the partials in k_dot's order, so r.r is the same as k_dot's)
__global__ __launch_bounds__(kRT) void k_mixed_start(const double* __restrict__ r64, double s,
                                                     float* __restrict__ r32, float* __restrict__ p32,
                                                     float* __restrict__ x32, int64_t n,
                                                     double* partials) {
  __shared__ double red[kRT];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const float r = (float)(r64[i] * -s);
    r32[i] = r;
    p32[i] = -r;
    x32[i] = 0.0f;
    acc += (double)r * (double)r;
  }
  const double t = block_reduce<kRT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// true residual: r = b - Ax (Ax in r on entry) and the partials of r.r
__global__ __launch_bounds__(kRT) void k_residual(const double* __restrict__ b,
                                                  double* __restrict__ r, int64_t n,
                                                  double* partials) {
  __shared__ double red[kRT];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const double v = b[i] - r[i];
    r[i] = v;
    acc += v * v;
  }
  const double t = block_reduce<kRT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// x += scale d; with p: d + alpha p first (alpha = scal[SC_ALPHA], in FP32
// as the CG x update computes it), the x update left by the inner loop
__global__ __launch_bounds__(kRT) void k_add_scaled(double* __restrict__ x,
                                                    const float* __restrict__ d, double scale,
                                                    int64_t n, const float* __restrict__ p,
                                                    const double* __restrict__ scal) {
  const int64_t stride = (int64_t)gridDim.x * kRT;
  const float al = p ? (float)scal[SC_ALPHA] : 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    float di = d[i];
    if (p) di += al * p[i];
    x[i] += scale * (double)di;
  }
}

struct Uniform {
  double v[9];
};

__global__ __launch_bounds__(kRT) void k_fill(double* __restrict__ F, Uniform u, int nc,
                                              int64_t npix, int add) {
  const int64_t n = npix * nc;
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const double v = u.v[i % nc];
    F[i] = add ? F[i] + v : v;
  }
}

// grid of the grid-stride kernels: at most `cap` workgroups (kRedGrid,
// measured best for kernels that read 16 B or more per value; the FP32 CG
// update without x reads 8 B per value and runs best with kRedBlocks)
inline unsigned grid_for(int64_t n, int cap = kRedGrid) {
  int64_t g = (n + kRT - 1) / kRT;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

__global__ void k_scal_copy(double* scal, int from, int to) {
  if (threadIdx.x == 0) scal[to] = scal[from];
}

// control words of a device-resident CG loop and its r.r threshold
__global__ void k_set_ctl(int* ctl, int maxit, double* scal, double tol2) {
  if (threadIdx.x == 0) {
    ctl[CTL_ACTIVE] = 1;
    ctl[CTL_ITER] = 0;
    ctl[CTL_MAXIT] = maxit;
    scal[SC_TOL2] = tol2;
  }
}

__global__ __launch_bounds__(kRT) void k_copy(const double2* __restrict__ s,
                                              double2* __restrict__ d, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) d[i] = s[i];
}

// Streaming copy for the bandwidth reference: every lane keeps kWide 16-byte
// loads in flight (all issued before the first store, from clamped
// addresses so that none waits behind a branch), optionally with
// non-temporal hints on the loads and stores.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kWide = 4;
template <bool NT>
__global__ __launch_bounds__(kRT) void k_copy_wide(const u32x4* __restrict__ s,
                                                   u32x4* __restrict__ d, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kRT * kWide;
  for (int64_t i0 = (int64_t)blockIdx.x * kRT * kWide + threadIdx.x; i0 < n; i0 += stride) {
    u32x4 v[kWide];
#pragma unroll
    for (int k = 0; k < kWide; ++k) {
      const int64_t i = i0 + (int64_t)k * kRT < n ? i0 + (int64_t)k * kRT : n - 1;
      v[k] = NT ? __builtin_nontemporal_load(s + i) : s[i];
    }
#pragma unroll
    for (int k = 0; k < kWide; ++k) {
      const int64_t i = i0 + (int64_t)k * kRT;
      if (i < n) {
        if (NT) __builtin_nontemporal_store(v[k], d + i);
        else d[i] = v[k];
      }
    }
  }
}

// ---- batched inner CG: one RVE per blockIdx (finalize) or blockIdx.z ----
// same arithmetic and reduction order as k_finalize / k_cg_update, so a slot
// computes exactly what the unbatched FP32 inner CG computes
__global__ __launch_bounds__(1024) void k_finalize_batch(const BatchSlot* __restrict__ slots,
                                                         int64_t count, int mode) {
  const BatchSlot& sl = slots[blockIdx.x];
  if (!sl.ctl[CTL_ACTIVE]) return;
  __shared__ double red[1024];
  const double* partials = mode == FIN_ALPHA ? sl.zpart : sl.upart;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += 1024) acc += partials[i];
  const double s = block_reduce<1024>(acc, red);
  if (threadIdx.x == 0) {
    double* scal = sl.scal;
    if (mode == FIN_ALPHA) {
      scal[SC_PAP] = s;
      scal[SC_ALPHA] = scal[SC_RDR] / s;
    } else {
      scal[SC_NEWRDR] = s;
      scal[SC_BETA] = s / scal[SC_RDR];
      scal[SC_RDR] = s;
      // loop test of the inner CG: iterations left and r.r above the
      // threshold; a NaN r.r stops the solve without counting the iteration
      if (!(s == s)) {
        sl.ctl[CTL_ACTIVE] = 0;
      } else {
        const int it = sl.ctl[CTL_ITER] + 1;
        sl.ctl[CTL_ITER] = it;
        if (!(it < sl.ctl[CTL_MAXIT] && s > scal[SC_TOL2])) sl.ctl[CTL_ACTIVE] = 0;
      }
    }
  }
}

// r += alpha Ap and the partials of r.r per slot; the x update rides on the
// next iteration's r2c (k_r2c_batch), and the last one is folded into the
// refinement's correction of x (add_scaled_convert with p)
__global__ __launch_bounds__(kRT) void k_cg_update_batch(const BatchSlot* __restrict__ slots,
                                                         int64_t n) {
  const BatchSlot& sl = slots[blockIdx.z];
  if (!sl.ctl[CTL_ACTIVE]) return;
  __shared__ double red[kRT];
  float* __restrict__ r = sl.r;
  const float* __restrict__ Ap = sl.Ap;
  const double alpha = sl.scal[SC_ALPHA];
  const float al = (float)alpha;
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kRT;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const float rv = r[i] + al * Ap[i];
    r[i] = rv;
    acc += (double)rv * (double)rv;
  }
  const double t = block_reduce<kRT>(acc, red);
  if (threadIdx.x == 0) sl.upart[blockIdx.x] = t;
}

}  // namespace

int batch_finalize(const BatchSlot* d_slots, int nslots, int64_t count, int mode) {
  if (nslots <= 0) return MUGPU_OK;
  hipLaunchKernelGGL(k_finalize_batch, dim3((unsigned)nslots), dim3(1024), 0, stream(), d_slots,
                     count, mode);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int batch_cg_update(const BatchSlot* d_slots, int nslots, int64_t n) {
  if (nslots <= 0) return MUGPU_OK;
  // the grid of the unbatched FP32 update (cg_update without x), so that
  // r.r is summed in the same order
  const unsigned g = grid_for(n, kRedBlocks);
  timing_begin(MUGPU_K_CG_UPDATE);
  hipLaunchKernelGGL(k_cg_update_batch, dim3(g, 1, (unsigned)nslots), dim3(kRT), 0, stream(),
                     d_slots, n);
  timing_end(MUGPU_K_CG_UPDATE);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int batch_cg_iteration(FftPlan& p, const BatchSlot* d_slots, int nslots) {
  const int64_t n = p.N * p.nc;
  MUGPU_TRY(fast_batch_operator(p, d_slots, nslots));
  MUGPU_TRY(batch_finalize(d_slots, nslots, p.nzb[1], FIN_ALPHA));
  MUGPU_TRY(batch_cg_update(d_slots, nslots, n));
  return batch_finalize(d_slots, nslots, grid_for(n, kRedBlocks), FIN_BETA);
}


int copy_device(void* dst, const void* src, size_t bytes) {
  const int64_t n = (int64_t)(bytes / 16);
  int64_t g = (n + kRT - 1) / kRT;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)g), dim3(kRT), 0, stream(), (const double2*)src,
                     (double2*)dst, n);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

// read-only and write-only streams for the bandwidth reference: a sum of
// every 16-byte value (kept live through one store per workgroup), and a fill
__global__ __launch_bounds__(kRT) void k_read_wide(const u32x4* __restrict__ s, int64_t n,
                                                   unsigned* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * kRT * kWide;
  unsigned acc = 0;
  for (int64_t i0 = (int64_t)blockIdx.x * kRT * kWide + threadIdx.x; i0 < n; i0 += stride) {
    u32x4 v[kWide];
#pragma unroll
    for (int k = 0; k < kWide; ++k) {
      const int64_t i = i0 + (int64_t)k * kRT < n ? i0 + (int64_t)k * kRT : n - 1;
      v[k] = __builtin_nontemporal_load(s + i);
    }
#pragma unroll
    for (int k = 0; k < kWide; ++k) acc += v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
  }
  if (acc == 0x9e3779b9u) sink[blockIdx.x] = acc;  // practically never taken
}

__global__ __launch_bounds__(kRT) void k_fill_wide(u32x4* __restrict__ d, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kRT;
  u32x4 z = {0u, 0u, 0u, 0u};
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(z, d + i);
}

int read_device_wide(const void* src, size_t bytes, void* sink, int grid) {
  const int64_t n = (int64_t)(bytes / 16);
  if (n < 1) return MUGPU_OK;
  const int64_t need = (n + (int64_t)kRT * kWide - 1) / ((int64_t)kRT * kWide);
  const unsigned g = (unsigned)(need < grid ? need : grid);
  hipLaunchKernelGGL(k_read_wide, dim3(g), dim3(kRT), 0, stream(), (const u32x4*)src, n,
                     (unsigned*)sink);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int fill_device_wide(void* dst, size_t bytes, int grid) {
  const int64_t n = (int64_t)(bytes / 16);
  if (n < 1) return MUGPU_OK;
  hipLaunchKernelGGL(k_fill_wide, dim3((unsigned)grid), dim3(kRT), 0, stream(), (u32x4*)dst, n);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int copy_device_wide(void* dst, const void* src, size_t bytes, bool nontemporal, int grid) {
  const int64_t n = (int64_t)(bytes / 16);
  if (n < 1) return MUGPU_OK;
  const int64_t need = (n + (int64_t)kRT * kWide - 1) / ((int64_t)kRT * kWide);
  const unsigned g = (unsigned)(need < grid ? need : grid);
  if (nontemporal)
    hipLaunchKernelGGL(k_copy_wide<true>, dim3(g), dim3(kRT), 0, stream(), (const u32x4*)src,
                       (u32x4*)dst, n);
  else
    hipLaunchKernelGGL(k_copy_wide<false>, dim3(g), dim3(kRT), 0, stream(), (const u32x4*)src,
                       (u32x4*)dst, n);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int dot_device(int prec, const void* a, const void* b, int64_t n, double* d_partials,
               double* d_scal, int mode) {
  const unsigned g = grid_for(n);
  if (prec == MUGPU_FP64)
    hipLaunchKernelGGL((k_dot<double, double>), dim3(g), dim3(kRT), 0, stream(),
                       (const double*)a, (const double*)b, n, d_partials);
  else
    hipLaunchKernelGGL((k_dot<float, float>), dim3(g), dim3(kRT), 0, stream(), (const float*)a,
                       (const float*)b, n, d_partials);
  MUGPU_HIP(hipGetLastError());
  return finalize(d_partials, g, d_scal, mode);
}

int set_ctl(int* ctl, int maxit, double* d_scal, double tol2) {
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(64), 0, stream(), ctl, maxit, d_scal, tol2);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int finalize(const double* d_partials, int64_t count, double* d_scal, int mode, int* ctl) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream(), d_partials, count, d_scal, mode,
                     ctl);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int cg_update(int prec, void* x, void* r, const void* p, const void* Ap, const double* d_scal,
              int64_t n, double* d_partials, int* ctl) {
  const unsigned g = grid_for(n, prec == MUGPU_FP32 && !x ? kRedBlocks : kRedGrid);
  timing_begin(MUGPU_K_CG_UPDATE);
  if (prec == MUGPU_FP64)
    hipLaunchKernelGGL((k_cg_update<double>), dim3(g), dim3(kRT), 0, stream(), (double*)x,
                       (double*)r, (const double*)p, (const double*)Ap, d_scal, n, d_partials, ctl);
  else
    hipLaunchKernelGGL((k_cg_update<float>), dim3(g), dim3(kRT), 0, stream(), (float*)x,
                       (float*)r, (const float*)p, (const float*)Ap, d_scal, n, d_partials, ctl);
  timing_end(MUGPU_K_CG_UPDATE);
  MUGPU_HIP(hipGetLastError());
  return finalize(d_partials, g, const_cast<double*>(d_scal), FIN_BETA, ctl);
}

int axpy_alpha(int prec, void* x, const void* p, const double* d_scal, int64_t n, int slot) {
  const unsigned g = grid_for(n);
  if (prec == MUGPU_FP64)
    hipLaunchKernelGGL((k_axpy_alpha<double>), dim3(g), dim3(kRT), 0, stream(), (double*)x,
                       (const double*)p, d_scal, n, slot);
  else
    hipLaunchKernelGGL((k_axpy_alpha<float>), dim3(g), dim3(kRT), 0, stream(), (float*)x,
                       (const float*)p, d_scal, n, slot);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int scal_copy(double* d_scal, int from, int to) {
  if (from < 0 || from >= SC_COUNT || to < 0 || to >= SC_COUNT) return MUGPU_EARG;
  hipLaunchKernelGGL(k_scal_copy, dim3(1), dim3(64), 0, stream(), d_scal, from, to);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int cg_init_vectors(int prec, const void* Ax, const void* b, void* r, void* p, int64_t n) {
  const unsigned g = grid_for(n);
  if (prec == MUGPU_FP64)
    hipLaunchKernelGGL((k_cg_init<double>), dim3(g), dim3(kRT), 0, stream(), (const double*)Ax,
                       (const double*)b, (double*)r, (double*)p, n);
  else
    hipLaunchKernelGGL((k_cg_init<float>), dim3(g), dim3(kRT), 0, stream(), (const float*)Ax,
                       (const float*)b, (float*)r, (float*)p, n);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int axpby(int prec, double a, const void* x, double b, void* y, int64_t n) {
  const unsigned g = grid_for(n);
  if (prec == MUGPU_FP64)
    hipLaunchKernelGGL((k_axpby<double>), dim3(g), dim3(kRT), 0, stream(), a, (const double*)x, b,
                       (double*)y, n);
  else
    hipLaunchKernelGGL((k_axpby<float>), dim3(g), dim3(kRT), 0, stream(), a, (const float*)x, b,
                       (float*)y, n);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int convert(int from_prec, const void* src, int to_prec, void* dst, int64_t n, double scale) {
  const unsigned g = grid_for(n);
  if (from_prec == MUGPU_FP64 && to_prec == MUGPU_FP32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(kRT), 0, stream(),
                       (const double*)src, (float*)dst, n, scale);
  else if (from_prec == MUGPU_FP32 && to_prec == MUGPU_FP64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(kRT), 0, stream(),
                       (const float*)src, (double*)dst, n, scale);
  else if (from_prec == MUGPU_FP64)
    hipLaunchKernelGGL((k_convert<double, double>), dim3(g), dim3(kRT), 0, stream(),
                       (const double*)src, (double*)dst, n, scale);
  else
    hipLaunchKernelGGL((k_convert<float, float>), dim3(g), dim3(kRT), 0, stream(),
                       (const float*)src, (float*)dst, n, scale);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int mixed_start(const double* r64, double s, float* r32, float* p32, float* x32, int64_t n,
                double* d_partials, double* d_scal, int mode) {
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_mixed_start, dim3(g), dim3(kRT), 0, stream(), r64, s, r32, p32, x32, n,
                     d_partials);
  MUGPU_HIP(hipGetLastError());
  return finalize(d_partials, g, d_scal, mode);
}

int residual_dot(const double* b, double* r, int64_t n, double* d_partials, double* d_scal,
                 int mode) {
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_residual, dim3(g), dim3(kRT), 0, stream(), b, r, n, d_partials);
  MUGPU_HIP(hipGetLastError());
  return finalize(d_partials, g, d_scal, mode);
}

int add_scaled_convert(void* x64, const void* d32, double scale, int64_t n, const void* p32,
                       const double* d_scal) {
  if (p32 && !d_scal) return MUGPU_EARG;
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_add_scaled, dim3(g), dim3(kRT), 0, stream(), (double*)x64,
                     (const float*)d32, scale, n, (const float*)p32, d_scal);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

int fill_uniform(double* F, const double* h_val, int nc, int64_t npix, int add) {
  Uniform u{};
  for (int c = 0; c < nc && c < 9; ++c) u.v[c] = h_val[c];
  const unsigned g = grid_for(npix * nc);
  hipLaunchKernelGGL(k_fill, dim3(g), dim3(kRT), 0, stream(), F, u, nc, npix, add);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

}  // namespace mugpu
