// Device-side building blocks of the gfx950 FFT passes.
//
// A workgroup holds `nl` complex lines of length n in LDS (line l, element m
// at s[l*ls + m]) and transforms all of them with mixed-radix Stockham
// autosort passes.  Each pass is register-staged so one LDS buffer suffices:
// every thread reads the inputs of its butterflies, the workgroup
// synchronises, and the thread writes the outputs in autosorted order.
//
//   pass with radix R and span Ns, for butterfly j < n/R, k = j mod Ns:
//     v_t = x[j + t n/R] * W_{Ns R}^{t k}      (t < R)
//     v   = DFT_R(v)
//     y[(j/Ns) Ns R + k + t Ns] = v_t
//
// Twiddles come from a per-axis table W_n^m = exp(-2 pi i m/n) (forward
// sign; the backward transform conjugates).  Radices 2, 3, 4, 5 and 8 have
// unrolled butterflies; any other prime factor uses a per-output direct sum.
#pragma once
#include <hip/hip_runtime.h>

namespace mugpu {

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };
template <typename T> using c_t = typename cplx<T>::type;

constexpr int kMaxFactors = 24;
constexpr int kEPT = 16;  // complex elements per thread held in registers

struct AxisPlan {
  int n;
  int nf;
  int f[kMaxFactors];
};

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  C r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  C r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
// multiply by -i (forward) or +i (backward)
template <typename C>
__device__ __forceinline__ C mul_mi(C a, int sign) {
  C r;
  if (sign < 0) {
    r.x = a.y;
    r.y = -a.x;
  } else {
    r.x = -a.y;
    r.y = a.x;
  }
  return r;
}
template <typename C>
__device__ __forceinline__ C twid(const C* tw, int e, int sign) {
  C w = tw[e];
  if (sign > 0) w.y = -w.y;
  return w;
}

// ---- butterflies --------------------------------------------------------
template <typename T, int R>
struct Butterfly;

template <typename T>
struct Butterfly<T, 2> {
  __device__ static void run(c_t<T>* v, int, const c_t<T>*, int) {
    c_t<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T>
struct Butterfly<T, 4> {
  __device__ static void run(c_t<T>* v, int sign, const c_t<T>*, int) {
    c_t<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    c_t<T> a2 = cadd(v[1], v[3]), a3 = mul_mi(csub(v[1], v[3]), sign);
    v[0] = cadd(a0, a2);
    v[1] = cadd(a1, a3);
    v[2] = csub(a0, a2);
    v[3] = csub(a1, a3);
  }
};

template <typename T>
struct Butterfly<T, 8> {
  // X[2q] = DFT4(v_t + v_{t+4}), X[2q+1] = DFT4((v_t - v_{t+4}) W8^t)
  __device__ static void run(c_t<T>* v, int sign, const c_t<T>*, int) {
    const T h = T(0.70710678118654752440084436210484903928);
    c_t<T> a[4], b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      a[t] = cadd(v[t], v[t + 4]);
      b[t] = csub(v[t], v[t + 4]);
    }
    // W8^1 = (1 -+ i)/sqrt2, W8^2 = -+i, W8^3 = (-1 -+ i)/sqrt2
    c_t<T> w1, w3;
    w1.x = h;
    w1.y = sign < 0 ? -h : h;
    w3.x = -h;
    w3.y = sign < 0 ? -h : h;
    b[1] = cmul(b[1], w1);
    b[2] = mul_mi(b[2], sign);
    b[3] = cmul(b[3], w3);
    Butterfly<T, 4>::run(a, sign, nullptr, 0);
    Butterfly<T, 4>::run(b, sign, nullptr, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[2 * q] = a[q];
      v[2 * q + 1] = b[q];
    }
  }
};

template <typename T>
struct Butterfly<T, 3> {
  __device__ static void run(c_t<T>* v, int sign, const c_t<T>*, int) {
    const T s3 = T(0.86602540378443864676372317075293618347);
    c_t<T> a = v[0], b = v[1], c = v[2];
    c_t<T> t1 = cadd(b, c);
    c_t<T> t2;
    t2.x = a.x - T(0.5) * t1.x;
    t2.y = a.y - T(0.5) * t1.y;
    c_t<T> d = csub(b, c);
    // i * s3 * d, sign-dependent
    c_t<T> t3;
    t3.x = -s3 * d.y;
    t3.y = s3 * d.x;
    v[0] = cadd(a, t1);
    if (sign < 0) {
      v[1] = csub(t2, t3);
      v[2] = cadd(t2, t3);
    } else {
      v[1] = cadd(t2, t3);
      v[2] = csub(t2, t3);
    }
  }
};

// radix 16 as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2,
//   X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 W4^(n1 k1) x[4 n1 + n2]
template <typename T>
struct Butterfly<T, 16> {
  __device__ static void run(c_t<T>* v, int sign, const c_t<T>*, int) {
    const T c1 = T(0.92387953251128675612818318939678828682);  // cos(pi/8)
    const T s1 = T(0.38268343236508977172845998403039886676);  // sin(pi/8)
    const T h = T(0.70710678118654752440084436210484903928);
    // cos and sin of 2 pi m / 16 for m = 0..9
    const T cs[10] = {T(1), c1, h, s1, T(0), -s1, -h, -c1, T(-1), -c1};
    const T sn[10] = {T(0), s1, h, c1, T(1), c1, h, s1, T(0), -s1};
    c_t<T> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      c_t<T> x[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      Butterfly<T, 4>::run(x, sign, nullptr, 0);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) {
        const int m = n2 * k1;
        if (m == 0) {
          a[n2][k1] = x[k1];
        } else if (m == 4) {
          a[n2][k1] = mul_mi(x[k1], sign);
        } else {
          c_t<T> w;
          w.x = cs[m];
          w.y = sign < 0 ? -sn[m] : sn[m];
          a[n2][k1] = cmul(x[k1], w);
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      c_t<T> y[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Butterfly<T, 4>::run(y, sign, nullptr, 0);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = y[k2];
    }
  }
};

// small odd radix by direct sums with roots W_R^m = tw[m * n/R]
template <typename T, int R>
struct Butterfly {
  __device__ static void run(c_t<T>* v, int sign, const c_t<T>* tw, int n) {
    c_t<T> u[R];
    const int step = n / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      c_t<T> acc = v[0];
#pragma unroll
      for (int t = 1; t < R; ++t) {
        c_t<T> w = twid(tw, ((t * q) % R) * step, sign);
        acc = cadd(acc, cmul(v[t], w));
      }
      u[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = u[q];
  }
};

// ---- one Stockham pass of radix R (register staged) ----------------------
template <typename T, int R, int NT>
__device__ __forceinline__ void pass_radix(c_t<T>* s, int nl, int ls, int n,
                                           int Ns, const c_t<T>* tw,
                                           int sign) {
  constexpr int MAXB = (kEPT + R - 1) / R;
  const int nb = n / R;
  const int total = nl * nb;
  c_t<T> v[MAXB][R];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = threadIdx.x + i * NT;
    if (b < total) {
      const int line = b / nb;
      const int j = b - line * nb;
      const c_t<T>* src = s + line * ls + j;
#pragma unroll
      for (int t = 0; t < R; ++t) v[i][t] = src[t * nb];
    }
  }
  __syncthreads();
  const int twstep = n / (Ns * R);
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = threadIdx.x + i * NT;
    if (b < total) {
      const int line = b / nb;
      const int j = b - line * nb;
      const int k = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t) v[i][t] = cmul(v[i][t], twid(tw, t * k * twstep, sign));
      }
      Butterfly<T, R>::run(v[i], sign, tw, n);
      c_t<T>* dst = s + line * ls + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) dst[t * Ns] = v[i][t];
    }
  }
  __syncthreads();
}

// generic prime radix: each output y[o] = sum_t x[j + t n/R] W_{Ns R}^{t (k + q Ns)}
template <typename T, int NT>
__device__ void pass_generic(c_t<T>* s, int nl, int ls, int n, int Ns, int R,
                             const c_t<T>* tw, int sign) {
  const int total = nl * n;
  const int nb = n / R;
  const int span = Ns * R;
  const int step = n / span;
  c_t<T> acc[kEPT];
#pragma unroll
  for (int i = 0; i < kEPT; ++i) {
    const int o = threadIdx.x + i * NT;
    if (o < total) {
      const int line = o / n;
      const int oo = o - line * n;
      const int k = oo % Ns;
      const int q = (oo / Ns) % R;
      const int j = (oo / span) * Ns + k;
      const int e = k + q * Ns;
      const c_t<T>* src = s + line * ls + j;
      c_t<T> a = src[0];
      // ex = t e mod span, kept by adding e (e < span) instead of a 64-bit
      // multiply and modulo per term
      int ex = 0;
      for (int t = 1; t < R; ++t) {
        ex += e;
        if (ex >= span) ex -= span;
        a = cadd(a, cmul(src[t * nb], twid(tw, ex * step, sign)));
      }
      acc[i] = a;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kEPT; ++i) {
    const int o = threadIdx.x + i * NT;
    if (o < total) {
      const int line = o / n;
      const int oo = o - line * n;
      s[line * ls + oo] = acc[i];
    }
  }
  __syncthreads();
}

// full transform of nl lines held in LDS; requires nl*n <= NT*kEPT
template <typename T, int NT>
__device__ void lds_fft(c_t<T>* s, int nl, int ls, const AxisPlan& ap,
                        const c_t<T>* tw, int sign) {
  int Ns = 1;
  const int n = ap.n;
  for (int f = 0; f < ap.nf; ++f) {
    const int R = ap.f[f];
    switch (R) {
      case 2: pass_radix<T, 2, NT>(s, nl, ls, n, Ns, tw, sign); break;
      case 3: pass_radix<T, 3, NT>(s, nl, ls, n, Ns, tw, sign); break;
      case 4: pass_radix<T, 4, NT>(s, nl, ls, n, Ns, tw, sign); break;
      case 5: pass_radix<T, 5, NT>(s, nl, ls, n, Ns, tw, sign); break;
      case 8: pass_radix<T, 8, NT>(s, nl, ls, n, Ns, tw, sign); break;
      default: pass_generic<T, NT>(s, nl, ls, n, Ns, R, tw, sign); break;
    }
    Ns *= R;
  }
}

}  // namespace mugpu
