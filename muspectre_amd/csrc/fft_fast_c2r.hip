// Last-axis c2r pass (c2r, p.Ap partials): launchers and dispatch on length
// and variant.
#include "fft_fast.hpp"

namespace mugpu {

namespace {

template <typename T, int N, int E, int DIM, int H>
int c2r_n(FftPlan& p, const OpSpec& op, void* spec) {
  LArgs<T> a{};
  a.spec = reinterpret_cast<c_t<T>*>(spec);
  a.out = reinterpret_cast<T*>(op.out);
  a.dotv = reinterpret_cast<const T*>(op.dotv);
  a.partials = reinterpret_cast<double*>(p.partials);
  a.rows = p.N / N;
  a.tw = tw_of<T>(p, DIM - 1);
  a.active = p.active;
  constexpr int RB = z_rows(N, DIM, H), NC = DIM * DIM;
  constexpr int NT = RB / 2 * NC * (N / E);
  constexpr size_t lds = TwLds<T, N>::total(lds_c2r<T, N, E, DIM, H>());
  const dim3 grid((unsigned)((a.rows + RB - 1) / RB));
  if (op.out_mode == OUT_DOT)
    hipLaunchKernelGGL((k_c2r_fast<T, N, E, DIM, OUT_DOT, H>), grid, dim3(NT), lds, stream(), a);
  else
    hipLaunchKernelGGL((k_c2r_fast<T, N, E, DIM, OUT_PLAIN, H>), grid, dim3(NT), lds, stream(), a);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

// last-axis passes: bit 0 of the variant selects E = 8, bit 1 half the rows
// per workgroup; a variant that changes nothing at this length runs the
// kernel it equals
template <typename T, int N, int E, int DIM>
int c2r_h(FftPlan& p, const OpSpec& op, void* spec, int h) {
  if constexpr (z_rows(N, DIM, 1) != z_rows(N, DIM, 0))
    if (h) return c2r_n<T, N, E, DIM, 1>(p, op, spec);
  return c2r_n<T, N, E, DIM, 0>(p, op, spec);
}
template <typename T, int DIM>
int c2r_any(FftPlan& p, const OpSpec& op, void* spec) {
  const int v = p.fvar[pi_of<T>()][FV_C2R][0];
#define C_(n, e) c2r_h<T, n, e, DIM>(p, op, spec, v >> 1)
  MUGPU_FAST_SWITCH(p.res[DIM - 1], v & 1, C_)
#undef C_
  return MUGPU_EARG;
}

}  // namespace

int fast_c2r(FftPlan& p, int prec, const OpSpec& op, void* spec) {
  if (prec == MUGPU_FP64) return p.dim == 3 ? c2r_any<double, 3>(p, op, spec) : c2r_any<double, 2>(p, op, spec);
  return p.dim == 3 ? c2r_any<float, 3>(p, op, spec) : c2r_any<float, 2>(p, op, spec);
}

}  // namespace mugpu
