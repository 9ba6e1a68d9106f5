// Strided c2c passes (middle axis, first axis of the engine transforms):
// launchers and dispatch on length and variant.
#include "fft_fast.hpp"

namespace mugpu {

namespace {

template <typename T, int N, int E>
int strided_n(FftPlan& p, int ax, void* spec, int sign) {
  int64_t inner = p.nc;
  for (int d = ax + 1; d < p.dim; ++d) inner *= p.hres[d];
  int64_t outer = 1;
  for (int d = 0; d < ax; ++d) outer *= p.hres[d];
  constexpr int L = s_lines<T, N, E>();
  constexpr int NT = L * (N / E);
  constexpr size_t lds = TwLds<T, N>::total(lds_strided<T, N, E>());
  const dim3 grid((unsigned)((inner + L - 1) / L), (unsigned)outer);
  c_t<T>* data = reinterpret_cast<c_t<T>*>(spec);
  if (sign < 0)
    hipLaunchKernelGGL((k_strided_fast<T, N, E, -1>), grid, dim3(NT), lds, stream(), data, inner, tw_of<T>(p, ax),
                       p.active);
  else
    hipLaunchKernelGGL((k_strided_fast<T, N, E, +1>), grid, dim3(NT), lds, stream(), data, inner, tw_of<T>(p, ax),
                       p.active);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <typename T>
int strided_any(FftPlan& p, int ax, void* spec, int sign) {
  const int v = p.fvar[pi_of<T>()][ax == 0 ? FV_STRIDED0 : FV_STRIDED1][0];
#define C_(n, e) strided_n<T, n, e>(p, ax, spec, sign)
  MUGPU_FAST_SWITCH(p.res[ax], v, C_)
#undef C_
  return MUGPU_EARG;
}
}  // namespace

int fast_strided(FftPlan& p, int prec, int ax, void* spec, int sign) {
  if (prec == MUGPU_FP64) return strided_any<double>(p, ax, spec, sign);
  return strided_any<float>(p, ax, spec, sign);
}

}  // namespace mugpu
