// First-axis pass with the projection: launchers and dispatch on length
// and variant.
#include "fft_fast.hpp"

namespace mugpu {

namespace {

template <typename T, int N, int E, int DIM, int H, bool B>
int axis0_n(FftPlan& p, const ProjParams& pj, void* spec) {
  XArgs<T> a{};
  a.data = reinterpret_cast<c_t<T>*>(spec);
  a.plane = p.kpts / p.res[0];
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
  constexpr int KP = x_kpts(N, DIM, sizeof(T) == 4, E, H), NC = DIM * DIM, L = KP * NC, NT = L * (N / E);
  constexpr size_t lds = TwLds<T, N>::total(lds_axis0<T, N, E, DIM, H>());
  const dim3 grid((unsigned)((a.plane + KP - 1) / KP));
  if constexpr (B)
    hipLaunchKernelGGL((k_axis0_bounded<T, N, E, DIM, H>), grid, dim3(NT), lds, stream(), a);
  else
    hipLaunchKernelGGL((k_axis0_fast<T, N, E, DIM, H>), grid, dim3(NT), lds, stream(), a);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

// first-axis pass: bit 0 of the variant selects E = 8, bit 1 half the
// k-points per workgroup, bit 2 the kernel with the occupancy hint; a
// variant that changes nothing at this length runs the kernel it equals
template <typename T, int N, int E, int DIM, bool B>
int axis0_h(FftPlan& p, const ProjParams& pj, void* spec, int h) {
  if constexpr (x_kpts(N, DIM, sizeof(T) == 4, E, 1) != x_kpts(N, DIM, sizeof(T) == 4, E, 0))
    if (h) return axis0_n<T, N, E, DIM, 1, B>(p, pj, spec);
  return axis0_n<T, N, E, DIM, 0, B>(p, pj, spec);
}
template <typename T, int N, int E, int DIM>
int axis0_hb(FftPlan& p, const ProjParams& pj, void* spec, int v) {
  if (v & 4) return axis0_h<T, N, E, DIM, true>(p, pj, spec, (v >> 1) & 1);
  return axis0_h<T, N, E, DIM, false>(p, pj, spec, (v >> 1) & 1);
}
template <typename T, int DIM>
int axis0_any(FftPlan& p, const ProjParams& pj, void* spec) {
  const int v = p.fvar[pi_of<T>()][FV_AXIS0][0];
#define C_(n, e) axis0_hb<T, n, e, DIM>(p, pj, spec, v)
  MUGPU_FAST_SWITCH(p.res[0], v & 1, C_)
#undef C_
  return MUGPU_EARG;
}

}  // namespace

int fast_axis0_project(FftPlan& p, int prec, const ProjParams& pj, void* spec) {
  if (prec == MUGPU_FP64) return p.dim == 3 ? axis0_any<double, 3>(p, pj, spec) : axis0_any<double, 2>(p, pj, spec);
  return p.dim == 3 ? axis0_any<float, 3>(p, pj, spec) : axis0_any<float, 2>(p, pj, spec);
}


}  // namespace mugpu
