// Batched operator of the serving loop (FP32, one RVE per blockIdx.z).
#include "fft_fast.hpp"

namespace mugpu {

namespace {

// ---- batched launchers (FP32, slots on blockIdx.z) ------------------------------
template <int N, int E, int DIM, int H>
int b_r2c(FftPlan& p, const BatchSlot* slots, int ns) {
  constexpr int RB = z_rows(N, DIM, H);
  constexpr int NT = RB / 2 * DIM * DIM * (N / E);
  const int64_t rows = p.N / N;
  constexpr size_t lds = TwLds<float, N>::total(lds_lines<float, N, DIM, H>());
  const dim3 grid((unsigned)((rows + RB - 1) / RB), 1, (unsigned)ns);
  hipLaunchKernelGGL((k_r2c_batch<N, E, DIM, H>), grid, dim3(NT), lds, stream(), slots, rows,
                     tw_of<float>(p, DIM - 1));
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <int N, int E, int DIM, int H>
int b_c2r(FftPlan& p, const BatchSlot* slots, int ns) {
  constexpr int RB = z_rows(N, DIM, H), NC = DIM * DIM;
  constexpr int NT = RB / 2 * NC * (N / E);
  const int64_t rows = p.N / N;
  constexpr size_t lds = TwLds<float, N>::total(lds_c2r<float, N, E, DIM, H>());
  const dim3 grid((unsigned)((rows + RB - 1) / RB), 1, (unsigned)ns);
  hipLaunchKernelGGL((k_c2r_batch<N, E, DIM, H>), grid, dim3(NT), lds, stream(), slots, rows,
                     tw_of<float>(p, DIM - 1));
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <int N, int E>
int b_strided(FftPlan& p, int ax, int sign, const BatchSlot* slots, int ns) {
  int64_t inner = p.nc;
  for (int d = ax + 1; d < p.dim; ++d) inner *= p.hres[d];
  int64_t outer = 1;
  for (int d = 0; d < ax; ++d) outer *= p.hres[d];
  constexpr int L = s_lines<float, N, E>();
  constexpr int NT = L * (N / E);
  constexpr size_t lds = TwLds<float, N>::total(lds_strided<float, N, E>());
  const dim3 grid((unsigned)((inner + L - 1) / L), (unsigned)outer, (unsigned)ns);
  if (sign < 0)
    hipLaunchKernelGGL((k_strided_batch<N, E, -1>), grid, dim3(NT), lds, stream(), slots, inner,
                       tw_of<float>(p, ax));
  else
    hipLaunchKernelGGL((k_strided_batch<N, E, +1>), grid, dim3(NT), lds, stream(), slots, inner,
                       tw_of<float>(p, ax));
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <int N, int E, int DIM, int H>
int b_axis0(FftPlan& p, const BatchSlot* slots, int ns) {
  XArgs<float> a{};
  a.plane = p.kpts / p.res[0];
  a.tw = tw_of<float>(p, 0);
  for (int d = 0; d < 3; ++d) a.res[d] = p.res[d];
  a.hlast = p.hres[DIM - 1];
  constexpr int KP = x_kpts(N, DIM, true, E, H), NC = DIM * DIM, L = KP * NC, NT = L * (N / E);
  constexpr size_t lds = TwLds<float, N>::total(lds_axis0<float, N, E, DIM, H>());
  const dim3 grid((unsigned)((a.plane + KP - 1) / KP), 1, (unsigned)ns);
  hipLaunchKernelGGL((k_axis0_batch<N, E, DIM, H>), grid, dim3(NT), lds, stream(), slots, a);
  MUGPU_HIP(hipGetLastError());
  return MUGPU_OK;
}

template <int N, int E, int DIM>
int b_r2c_h(FftPlan& p, const BatchSlot* slots, int ns, int h) {
  if constexpr (z_rows(N, DIM, 1) != z_rows(N, DIM, 0))
    if (h) return b_r2c<N, E, DIM, 1>(p, slots, ns);
  return b_r2c<N, E, DIM, 0>(p, slots, ns);
}
template <int N, int E, int DIM>
int b_c2r_h(FftPlan& p, const BatchSlot* slots, int ns, int h) {
  if constexpr (z_rows(N, DIM, 1) != z_rows(N, DIM, 0))
    if (h) return b_c2r<N, E, DIM, 1>(p, slots, ns);
  return b_c2r<N, E, DIM, 0>(p, slots, ns);
}
template <int N, int E, int DIM>
int b_axis0_h(FftPlan& p, const BatchSlot* slots, int ns, int h) {
  if constexpr (x_kpts(N, DIM, true, E, 1) != x_kpts(N, DIM, true, E, 0))
    if (h) return b_axis0<N, E, DIM, 1>(p, slots, ns);
  return b_axis0<N, E, DIM, 0>(p, slots, ns);
}

template <int DIM>
int batch_operator_dim(FftPlan& p, const BatchSlot* slots, int ns) {
  int rc = MUGPU_EARG;
  const unsigned char(&fv)[FV_COUNT][2] = p.fvar[1];
  timing_begin(MUGPU_K_LAST_R2C);
#define C_(n, e) b_r2c_h<n, e, DIM>(p, slots, ns, (fv[FV_R2C][1] >> 1) & 1)
  rc = [&]() -> int { MUGPU_FAST_SWITCH(p.res[DIM - 1], fv[FV_R2C][1] & 1, C_) return MUGPU_EARG; }();
#undef C_
  timing_end(MUGPU_K_LAST_R2C);
  if (rc) return rc;
  if (DIM == 3) {
    timing_begin(MUGPU_K_MIDDLE);
#define C_(n, e) b_strided<n, e>(p, 1, -1, slots, ns)
    rc = [&]() -> int { MUGPU_FAST_SWITCH(p.res[1], fv[FV_STRIDED1][0], C_) return MUGPU_EARG; }();
#undef C_
    timing_end(MUGPU_K_MIDDLE);
    if (rc) return rc;
  }
  timing_begin(MUGPU_K_AXIS0);
// (bit 2 of the first-axis variant, the occupancy hint, belongs to the
// single-cell kernel; the batched one takes bits 0 and 1)
#define C_(n, e) b_axis0_h<n, e, DIM>(p, slots, ns, (fv[FV_AXIS0][0] >> 1) & 1)
  rc = [&]() -> int { MUGPU_FAST_SWITCH(p.res[0], fv[FV_AXIS0][0] & 1, C_) return MUGPU_EARG; }();
#undef C_
  timing_end(MUGPU_K_AXIS0);
  if (rc) return rc;
  if (DIM == 3) {
    timing_begin(MUGPU_K_MIDDLE);
#define C_(n, e) b_strided<n, e>(p, 1, +1, slots, ns)
    rc = [&]() -> int { MUGPU_FAST_SWITCH(p.res[1], fv[FV_STRIDED1][0], C_) return MUGPU_EARG; }();
#undef C_
    timing_end(MUGPU_K_MIDDLE);
    if (rc) return rc;
  }
  timing_begin(MUGPU_K_LAST_C2R);
#define C_(n, e) b_c2r_h<n, e, DIM>(p, slots, ns, (fv[FV_C2R][0] >> 1) & 1)
  rc = [&]() -> int { MUGPU_FAST_SWITCH(p.res[DIM - 1], fv[FV_C2R][0] & 1, C_) return MUGPU_EARG; }();
#undef C_
  timing_end(MUGPU_K_LAST_C2R);
  return rc;
}

}  // namespace

int fast_batch_operator(FftPlan& p, const BatchSlot* d_slots, int nslots) {
  for (int d = 0; d < p.dim; ++d)
    if (!p.fastax[d]) {
      set_error("batched solves need axis lengths with compiled passes (16, 32, 48, 64, 96, 128, 256, 512), got %d",
                p.res[d]);
      return MUGPU_EARG;
    }
  if (nslots <= 0) return MUGPU_OK;
  return p.dim == 3 ? batch_operator_dim<3>(p, d_slots, nslots)
                    : batch_operator_dim<2>(p, d_slots, nslots);
}


}  // namespace mugpu
