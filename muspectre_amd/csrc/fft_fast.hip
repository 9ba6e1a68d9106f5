// Compile-time-length FFT passes: supported lengths, workgroup shapes
// exported to the plans, and the variant of each pass chosen by the
// autotuner.  The kernels are in fft_fast.hpp.
#include "fft_fast.hpp"

namespace mugpu {

bool fast_size(int n) {
  return n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 256 || n == 512;
}

int fast_rows_per_block(int n, int dim, int variant) { return z_rows(n, dim, (variant >> 1) & 1); }

// ---- HBM pattern of the first-axis pass, without the transforms ------------
// Workgroup b takes the columns of tile xcd_tile(b) (kp k-points of nc
// components, a run of kp * nc complex values on every one of the n rows) and
// thread (l, p) reads rows p + (n / 16) i, i < 16, of column l into
// registers, then writes them back scaled by f: the loads, stores, tile order
// and workgroup shape of k_axis0_fast with 16 values per thread, and nothing
// in between.
namespace {
template <typename T>
__global__ __launch_bounds__(1024) void k_line_pattern(c_t<T>* __restrict__ data, int n,
                                                       int64_t plane, int nc, int kp, T f) {
  constexpr int E = 16;
  const int L = kp * nc;
  const int l = threadIdx.x % L, p = threadIdx.x / L, step = n / E;
  const int64_t inner = plane * nc;
  const int64_t w = xcd_tile(blockIdx.x, gridDim.x) * kp * nc + l;
  if (w >= inner) return;
  c_t<T>* base = data + w;
  c_t<T> v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = base[(int64_t)(p + step * i) * inner];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    v[i].x *= f;
    v[i].y *= f;
    base[(int64_t)(p + step * i) * inner] = v[i];
  }
}
}  // namespace

int line_pattern(void* data, int prec, int n, int64_t plane, int nc, int kp) {
  if (n % 16 || kp < 1 || nc < 1 || (int64_t)kp * nc * (n / 16) > 1024) return MUGPU_EARG;
  const dim3 grid((unsigned)((plane + kp - 1) / kp)), block((unsigned)(kp * nc * (n / 16)));
  if (prec == MUGPU_FP64)
    hipLaunchKernelGGL(k_line_pattern<double>, grid, block, 0, stream(),
                       reinterpret_cast<double2*>(data), n, plane, nc, kp, 1.0);
  else
    hipLaunchKernelGGL(k_line_pattern<float>, grid, block, 0, stream(),
                       reinterpret_cast<float2*>(data), n, plane, nc, kp, 1.0f);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

// variants chosen by the autotuner (muspectre_amd/autotune.py writes this
// table from timings of every variant on the GPU); lengths with a factor 3
// have one variant
namespace {
struct FastTuned {
  int pass, n, dim, prec, iso, variant;
};
#include "fft_tuned.inc"
}  // namespace

// variants forced by mugpu_debug_force_variant, per precision (0 FP64,
// 1 FP32) and pass; -1: the tuned table
namespace {
int g_forced[2][FV_COUNT] = {{-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}};
}  // namespace

void force_variant(int pass, int prec, int variant) {
  g_forced[prec == MUGPU_FP64 ? 0 : 1][pass] = variant;
}

int fast_variant(int pass, int n, int dim, int prec, int iso) {
  const int forced = g_forced[prec == MUGPU_FP64 ? 0 : 1][pass];
  if (forced >= 0 && fast_variant_valid(pass, n, dim, prec, forced)) return forced;
  for (const FastTuned& t : kFastTuned)
    if (t.pass == pass && t.n == n && t.dim == dim && t.prec == prec && t.iso == iso &&
        fast_variant_valid(pass, n, dim, prec, t.variant))
      return t.variant;
  return 0;
}

bool fast_variant_valid(int pass, int n, int dim, int prec, int variant) {
  const bool halves = pass == FV_AXIS0 || pass == FV_R2C || pass == FV_C2R;
  // bit 2: the first-axis kernel with the occupancy hint (k_axis0_bounded)
  if (variant < 0 || variant > (pass == FV_AXIS0 ? 7 : halves ? 3 : 1)) return false;
  if ((variant & 1) && n % 3 == 0) return false;
  if (variant & 2) {
    if (pass != FV_AXIS0) return z_rows(n, dim, 1) != z_rows(n, dim, 0);
    const int e = (variant & 1) ? 8 : fe_default(n);
    const bool f32 = prec != MUGPU_FP64;
    return x_kpts(n, dim, f32, e, 1) != x_kpts(n, dim, f32, e, 0);
  }
  return true;
}

}  // namespace mugpu
