// Device stages fused around the FFT passes: the directional-stiffness
// product K : p before the forward transform, the Green-operator projection
// per wave vector between the transforms, and block reductions for the
// p . Ap partial sums after the inverse transform.
#pragma once
#include "mugpu_internal.hpp"

namespace mugpu {

// dp = K : p for voxel v in the cell's tangent representation (blocked
// layout of tan_off(); lanes on consecutive voxels read contiguous runs)
template <typename T, int DIM>
__device__ __forceinline__ void tangent_product(int kind, const T* __restrict__ tan,
                                                int64_t v, const T* p, T* dp) {
  constexpr int NC = DIM * DIM;
  if (kind == TAN_ISO) {
    const T lam = tan[tan_off(v, 0, 2)], mu = tan[tan_off(v, 1, 2)];
    T tr = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) tr += p[i * (DIM + 1)];
#pragma unroll
    for (int j = 0; j < DIM; ++j)
#pragma unroll
      for (int i = 0; i < DIM; ++i)
        dp[i + DIM * j] = (i == j ? lam * tr : T(0)) + mu * (p[i + DIM * j] + p[j + DIM * i]);
  } else if (kind == TAN_HEP) {
    constexpr int S = DIM * DIM * DIM;
    const T* M = tan + tan_off(v, 0, S);
    T s[DIM];
#pragma unroll
    for (int n = 0; n < DIM; ++n) {
      T a = 0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) a += p[j + DIM * n];
      s[n] = a;
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      T a = 0;
#pragma unroll
      for (int n = 0; n < DIM; ++n) a += M[(q + NC * n) * kTanBlock] * s[n];
      dp[q] = a;
    }
  } else if (kind == TAN_SYM) {
    constexpr int S = NC * (NC + 1) / 2;
    const T* K = tan + tan_off(v, 0, S);
#pragma unroll
    for (int r = 0; r < NC; ++r) dp[r] = T(0);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      const T kd = K[sym_index(r, r, NC) * kTanBlock];
      dp[r] += kd * p[r];
#pragma unroll
      for (int c = r + 1; c < NC; ++c) {
        const T k = K[sym_index(r, c, NC) * kTanBlock];
        dp[r] += k * p[c];
        dp[c] += k * p[r];
      }
    }
  } else {
    const T* K = tan + tan_off(v, 0, NC * NC);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      T a = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) a += K[(r + NC * c) * kTanBlock] * p[c];
      dp[r] = a;
    }
  }
}

template <typename T>
__device__ __forceinline__ T fft_freq(int c, int n, double L) {
  // numpy fftfreq(n)*n order divided by the length (src/fft/fft_utils.cc:41-56)
  const int npos = (n - 1) / 2 + 1;
  return T((double)(c < npos ? c : c - n) / L);
}

// Green-operator projection of the nc complex components of one wave
// vector, f[(i + DIM j) * stride], in place, with unit wave vector xi and the
// normalisation 1/N folded in:
//   small strain (src/fft/projection_small_strain.cc:61-85):
//     G f = (xi_i a_j + xi_i b_j + xi_j a_i + xi_j b_i)/2 - xi_i xi_j s,
//     a = f xi, b = f^T xi, s = xi^T f xi
//   finite strain (projection_finite_strain_fast.cc:65-85 and the full-G
//   variant, projection_finite_strain.cc:61-95, which is the same operator):
//     f_ij <- (sum_k f_ik xi_k) xi_j
template <typename T, int DIM>
__device__ __forceinline__ void project_point(c_t<T>* f, int stride, const T* xi, int kind,
                                              T norm) {
  constexpr int NC = DIM * DIM;
  T fr[NC], fi[NC];
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    const c_t<T> v = f[r * stride];
    fr[r] = v.x;
    fi[r] = v.y;
  }
  if (kind == MUGPU_PROJ_SMALL_STRAIN) {
    T ar[DIM], br[DIM], ai[DIM], bi[DIM];
    T sr_ = 0, si_ = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      T a1 = 0, b1 = 0, a2 = 0, b2 = 0;
#pragma unroll
      for (int m = 0; m < DIM; ++m) {
        a1 += fr[i + DIM * m] * xi[m];
        a2 += fi[i + DIM * m] * xi[m];
        b1 += fr[m + DIM * i] * xi[m];
        b2 += fi[m + DIM * i] * xi[m];
      }
      ar[i] = a1;
      ai[i] = a2;
      br[i] = b1;
      bi[i] = b2;
      sr_ += xi[i] * a1;
      si_ += xi[i] * a2;
    }
#pragma unroll
    for (int j = 0; j < DIM; ++j)
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        const T gr = T(0.5) * (xi[i] * ar[j] + xi[i] * br[j] + xi[j] * ar[i] + xi[j] * br[i]) -
                     xi[i] * xi[j] * sr_;
        const T gi = T(0.5) * (xi[i] * ai[j] + xi[i] * bi[j] + xi[j] * ai[i] + xi[j] * bi[i]) -
                     xi[i] * xi[j] * si_;
        c_t<T> g;
        g.x = norm * gr;
        g.y = norm * gi;
        f[(i + DIM * j) * stride] = g;
      }
  } else {
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      T s1 = 0, s2 = 0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        s1 += fr[i + DIM * k] * xi[k];
        s2 += fi[i + DIM * k] * xi[k];
      }
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        c_t<T> g;
        g.x = norm * (s1 * xi[j]);
        g.y = norm * (s2 * xi[j]);
        f[(i + DIM * j) * stride] = g;
      }
    }
  }
}

// unit wave vector of Fourier point c (first DIM coordinates) of the grid
// res with lengths L; false for k = 0 and, under the even-grid policy, for a
// Nyquist index on an even axis (the operator is zero there).  |xi|^2 is
// summed without contraction into fused multiply-adds: left to the compiler,
// whether it contracts depends on the kernel the function is inlined into,
// and the serving and single-cell first-axis kernels then differ in the last
// bit of xi wherever 1/L is inexact (lengths like 7 or 5)
template <typename T, int DIM>
__device__ __forceinline__ bool unit_wave_vector(const int* c, const int* res, const double* L,
                                                 int allow_even, T* xi) {
#pragma clang fp contract(off)
  T n2 = 0;
  bool zero = true, nyq = false;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    xi[d] = fft_freq<T>(c[d], res[d], L[d]);
    n2 += xi[d] * xi[d];
    if (c[d] != 0) zero = false;
    if (allow_even && (res[d] % 2 == 0) && 2 * c[d] == res[d]) nyq = true;
  }
  if (zero || nyq) return false;
  const T inv = T(1) / sqrt(n2);
#pragma unroll
  for (int d = 0; d < DIM; ++d) xi[d] *= inv;
  return true;
}

// deterministic sum over the NT threads of a block (any NT <= 1024); red
// holds NT doubles; every thread must call it
template <int NT>
__device__ double block_sum_any(double v, double* red) {
  constexpr int P2 = NT <= 64 ? 64 : NT <= 128 ? 128 : NT <= 256 ? 256 : NT <= 512 ? 512 : 1024;
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = P2 / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s && threadIdx.x + s < NT) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

}  // namespace mugpu
