// Last-axis r2c pass (p update, K:p, r2c): launchers and dispatch on length
// and variant.
#include "fft_fast.hpp"

namespace mugpu {

namespace {

template <typename T, int N, int E, int DIM, int H>
int r2c_n(FftPlan& p, const OpSpec& op, void* spec) {
  LArgs<T> a{};
  a.in = reinterpret_cast<const T*>(op.in);
  a.rvec = reinterpret_cast<const T*>(op.rvec);
  a.pout = reinterpret_cast<T*>(op.pout);
  a.scal = op.scal;
  a.tan = reinterpret_cast<const T*>(op.tangent);
  a.tan_kind = op.tan_kind;
  a.spec = reinterpret_cast<c_t<T>*>(spec);
  a.rows = p.N / N;
  a.tw = tw_of<T>(p, DIM - 1);
  a.active = p.active;
  a.x = reinterpret_cast<T*>(op.xvec);
  constexpr int RB = z_rows(N, DIM, H);
  constexpr int NT = RB / 2 * DIM * DIM * (N / E);
  constexpr size_t lds = TwLds<T, N>::total(lds_lines<T, N, DIM, H>());
  const dim3 grid((unsigned)((a.rows + RB - 1) / RB));
  const bool iso = op.tan_kind == TAN_ISO;
  // the fused x update goes with the p update only
  const bool xu = op.xvec != nullptr;
  if (xu && op.in_mode != IN_TAN_PUPDATE) return MUGPU_EARG;
  if (op.in_mode == IN_PLAIN)
    hipLaunchKernelGGL((k_r2c_fast<T, N, E, DIM, IN_PLAIN, false, H>), grid, dim3(NT), lds, stream(), a);
  else if (op.in_mode == IN_TAN && iso)
    hipLaunchKernelGGL((k_r2c_fast<T, N, E, DIM, IN_TAN, true, H>), grid, dim3(NT), lds, stream(), a);
  else if (op.in_mode == IN_TAN)
    hipLaunchKernelGGL((k_r2c_fast<T, N, E, DIM, IN_TAN, false, H>), grid, dim3(NT), lds, stream(), a);
  else if (iso && xu)
    hipLaunchKernelGGL((k_r2c_fast<T, N, E, DIM, IN_TAN_PUPDATE_X, true, H>), grid, dim3(NT), lds, stream(), a);
  else if (iso)
    hipLaunchKernelGGL((k_r2c_fast<T, N, E, DIM, IN_TAN_PUPDATE, true, H>), grid, dim3(NT), lds, stream(), a);
  else if (xu)
    hipLaunchKernelGGL((k_r2c_fast<T, N, E, DIM, IN_TAN_PUPDATE_X, false, H>), grid, dim3(NT), lds, stream(), a);
  else
    hipLaunchKernelGGL((k_r2c_fast<T, N, E, DIM, IN_TAN_PUPDATE, false, H>), grid, dim3(NT), lds, stream(), a);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

// last-axis passes: bit 0 of the variant selects E = 8, bit 1 half the rows
// per workgroup; a variant that changes nothing at this length runs the
// kernel it equals
template <typename T, int N, int E, int DIM>
int r2c_h(FftPlan& p, const OpSpec& op, void* spec, int h) {
  if constexpr (z_rows(N, DIM, 1) != z_rows(N, DIM, 0))
    if (h) return r2c_n<T, N, E, DIM, 1>(p, op, spec);
  return r2c_n<T, N, E, DIM, 0>(p, op, spec);
}
template <typename T, int DIM>
int r2c_any(FftPlan& p, const OpSpec& op, void* spec) {
  const int v = p.fvar[pi_of<T>()][FV_R2C][op.tan_kind == TAN_ISO ? 1 : 0];
#define C_(n, e) r2c_h<T, n, e, DIM>(p, op, spec, v >> 1)
  MUGPU_FAST_SWITCH(p.res[DIM - 1], v & 1, C_)
#undef C_
  return MUGPU_EARG;
}

}  // namespace

int fast_r2c(FftPlan& p, int prec, const OpSpec& op, void* spec) {
  if (prec == MUGPU_FP64) return p.dim == 3 ? r2c_any<double, 3>(p, op, spec) : r2c_any<double, 2>(p, op, spec);
  return p.dim == 3 ? r2c_any<float, 3>(p, op, spec) : r2c_any<float, 2>(p, op, spec);
}

}  // namespace mugpu
