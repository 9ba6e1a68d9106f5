// Compile-time-length FFT passes: supported lengths, workgroup shapes
// exported to the plans, and the variant of each pass chosen by the
// autotuner.  The kernels are in fft_fast.hpp.
#include "fft_fast.hpp"

namespace mugpu {

bool fast_size(int n) {
  return n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 256 || n == 512;
}

int fast_rows_per_block(int n, int dim, int variant) { return z_rows(n, dim, (variant >> 1) & 1); }

// variants chosen by the autotuner (muspectre_amd/autotune.py writes this
// table from timings of every variant on the GPU); lengths with a factor 3
// have one variant
namespace {
struct FastTuned {
  int pass, n, dim, prec, iso, variant;
};
#include "fft_tuned.inc"
}  // namespace

int fast_variant(int pass, int n, int dim, int prec, int iso) {
  for (const FastTuned& t : kFastTuned)
    if (t.pass == pass && t.n == n && t.dim == dim && t.prec == prec && t.iso == iso &&
        fast_variant_valid(pass, n, dim, prec, t.variant))
      return t.variant;
  return 0;
}

bool fast_variant_valid(int pass, int n, int dim, int prec, int variant) {
  const bool halves = pass == FV_AXIS0 || pass == FV_R2C || pass == FV_C2R;
  if (variant < 0 || variant > (halves ? 3 : 1)) return false;
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
