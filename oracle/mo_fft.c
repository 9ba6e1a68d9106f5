/*
 * CPU FFT of the oracle: mixed-radix Stockham 1D transforms and the
 * multi-dimensional r2c / c2r over AoS fields with nc interleaved
 * components, as the reference's FFTW plans compute them
 * (src/fft/fftw_engine.cc:59-77 rank=d, howmany=nc, istride=ostride=nc,
 * idist=odist=1; r2c :98-100, c2r :108-110).  FFTW itself is an
 * un-vendored dependency with no pinned version (cmake/FindFFTW.cmake);
 * its published contract is the unnormalised DFT
 *     X[k] = sum_m x[m] exp(-2 pi i k.m / n)
 * with the last axis halved to n/2+1 Hermitian points.  Parity of these
 * routines is pinned against numpy.fft.rfftn/irfftn and a brute-force DFT
 * (tests/test_oracle_fft.py).
 *
 * Test infrastructure only.
 */
#include "muspectre_oracle.h"
#include "mo_internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;
static char g_err[512];

void mo_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int mo_get_threads(void) { return g_threads; }
const char *mo_last_error(void) { return g_err; }

void mo_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

long mo_grid_size(int dim, const int *res) {
  long n = 1;
  for (int i = 0; i < dim; ++i) n *= res[i];
  return n;
}

/* src/common/ccoord_operations.hh:63-68,101-104: last axis halved */
void mo_hermitian_sizes(int dim, const int *res, int *out) {
  for (int i = 0; i < dim; ++i) out[i] = res[i];
  out[dim - 1] = res[dim - 1] / 2 + 1;
}

/* src/fft/fft_utils.cc:41-56 */
void mo_fft_freqs(int n, double length, double *out) {
  int npos = (n - 1) / 2 + 1;
  for (int i = 0; i < npos; ++i) out[i] = (double)i / length;
  for (int i = npos; i < n; ++i)
    out[i] = (double)(-(n / 2) + i - npos) / length;
}

/* ---------------------------------------------------------------------- */
int mo_fplan_init(mo_fplan *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  /* radix 4 first, then 2, then odd primes in increasing order */
  while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
  for (int q = 3; m > 1; q += 2) {
    while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
  }
  p->tw = (mo_cpx *)malloc(sizeof(mo_cpx) * (size_t)(n > 0 ? n : 1));
  if (!p->tw) return MO_ALLOC_ERROR;
  for (int j = 0; j < n; ++j) {
    /* exp(-2 pi i j / n); angle reduced to [0, pi/4] by symmetry so
     * that the table is accurate to ~1 ulp */
    long jj = j, nn = n;
    /* use the octant symmetry on 8*j versus n */
    double c, s;
    long a = 8 * jj; /* angle = 2 pi j/n = (pi/4) * (8j/n) */
    long oct = a / nn;
    double frac = (double)(a - oct * nn) / (double)nn; /* in [0,1) */
    double t = frac * (M_PI / 4.0);
    double ct = cos(t), st = sin(t);
    double ctc = cos(M_PI / 4.0 - t), stc = sin(M_PI / 4.0 - t);
    switch (oct) {
      case 0: c = ct; s = st; break;
      case 1: c = stc; s = ctc; break;
      case 2: c = -st; s = ct; break;
      case 3: c = -ctc; s = stc; break;
      case 4: c = -ct; s = -st; break;
      case 5: c = -stc; s = -ctc; break;
      case 6: c = st; s = -ct; break;
      default: c = ctc; s = -stc; break;
    }
    p->tw[j].re = c;
    p->tw[j].im = -s;
  }
  return MO_OK;
}

void mo_fplan_free(mo_fplan *p) {
  free(p->tw);
  p->tw = NULL;
}

static inline mo_cpx cmul(mo_cpx a, mo_cpx b) {
  mo_cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}

/* W_n^e with the transform sign applied (sign=-1: table value, +1: conj) */
static inline mo_cpx twid(const mo_fplan *p, long e, int sign) {
  mo_cpx w = p->tw[e % p->n];
  if (sign > 0) w.im = -w.im;
  return w;
}

/*
 * Stockham autosort, mixed radix: pass with radix R and span Ns
 *   for j < n/R:  k = j mod Ns
 *     v_t = x[j + t n/R] * W_{Ns R}^{t k}
 *     v   = DFT_R(v)
 *     y[(j/Ns) Ns R + k + t Ns] = v_t
 * which leaves the output in natural order.
 */
void mo_fft_line(const mo_fplan *p, mo_cpx *x, mo_cpx *scratch, int sign) {
  const int n = p->n;
  if (n <= 1) return;
  mo_cpx *in = x, *out = scratch;
  long Ns = 1;
  mo_cpx v[64], u[64];
  for (int f = 0; f < p->nf; ++f) {
    const int R = p->f[f];
    const long stride = n / R;
    const long twstep = n / (Ns * R);
    for (long j = 0; j < stride; ++j) {
      const long k = j % Ns;
      const long base = (j / Ns) * Ns * R + k;
      if (R <= 64) {
        for (int t = 0; t < R; ++t) {
          mo_cpx a = in[j + t * stride];
          v[t] = t == 0 ? a : cmul(a, twid(p, (long)t * k * twstep, sign));
        }
        if (R == 2) {
          mo_cpx a = v[0], b = v[1];
          out[base].re = a.re + b.re; out[base].im = a.im + b.im;
          out[base + Ns].re = a.re - b.re; out[base + Ns].im = a.im - b.im;
        } else if (R == 4) {
          mo_cpx a0 = {v[0].re + v[2].re, v[0].im + v[2].im};
          mo_cpx a1 = {v[0].re - v[2].re, v[0].im - v[2].im};
          mo_cpx a2 = {v[1].re + v[3].re, v[1].im + v[3].im};
          mo_cpx a3 = {v[1].re - v[3].re, v[1].im - v[3].im};
          /* multiply a3 by W_4^{1} = -i (forward) or +i (backward) */
          mo_cpx a3r = sign < 0 ? (mo_cpx){a3.im, -a3.re}
                                : (mo_cpx){-a3.im, a3.re};
          out[base].re = a0.re + a2.re; out[base].im = a0.im + a2.im;
          out[base + Ns].re = a1.re + a3r.re; out[base + Ns].im = a1.im + a3r.im;
          out[base + 2 * Ns].re = a0.re - a2.re; out[base + 2 * Ns].im = a0.im - a2.im;
          out[base + 3 * Ns].re = a1.re - a3r.re; out[base + 3 * Ns].im = a1.im - a3r.im;
        } else {
          /* generic DFT_R: u_q = sum_t v_t W_R^{t q} */
          const long rstep = n / R;
          for (int q = 0; q < R; ++q) {
            mo_cpx acc = v[0];
            for (int t = 1; t < R; ++t) {
              mo_cpx w = twid(p, ((long)t * q % R) * rstep, sign);
              mo_cpx m = cmul(v[t], w);
              acc.re += m.re;
              acc.im += m.im;
            }
            u[q] = acc;
          }
          for (int q = 0; q < R; ++q) out[base + q * Ns] = u[q];
        }
      } else {
        /* very large prime radix: direct sums straight from memory */
        const long rstep = n / R;
        for (int q = 0; q < R; ++q) {
          mo_cpx acc = {0.0, 0.0};
          for (int t = 0; t < R; ++t) {
            mo_cpx a = in[j + t * stride];
            if (t) a = cmul(a, twid(p, (long)t * k * twstep, sign));
            mo_cpx m = cmul(a, twid(p, ((long)t * q % R) * rstep, sign));
            acc.re += m.re;
            acc.im += m.im;
          }
          out[base + q * Ns] = acc;
        }
      }
    }
    mo_cpx *tmp = in;
    in = out;
    out = tmp;
    Ns *= R;
  }
  if (in != x) memcpy(x, in, sizeof(mo_cpx) * (size_t)n);
}

/* ---------------------------------------------------------------------- */
/* c2c along axis `ax` of an AoS complex array with dims `h` (d dims) and nc
 * components; sign -1 forward, +1 backward */
static int c2c_axis(int d, const int *h, int nc, mo_cpx *data, int ax,
                    int sign) {
  const int n = h[ax];
  if (n <= 1) return MO_OK;
  long inner = nc; /* stride of axis ax, in complex elements */
  for (int a = ax + 1; a < d; ++a) inner *= h[a];
  long outer = 1;
  for (int a = 0; a < ax; ++a) outer *= h[a];
  mo_fplan p;
  if (mo_fplan_init(&p, n)) return MO_ALLOC_ERROR;
  const long nlines = outer * inner;
  int err = 0;
#pragma omp parallel num_threads(g_threads)
  {
    mo_cpx *buf = (mo_cpx *)malloc(sizeof(mo_cpx) * 2 * (size_t)n);
    if (!buf) {
#pragma omp atomic write
      err = 1;
    } else {
#pragma omp for schedule(static)
      for (long l = 0; l < nlines; ++l) {
        const long o = l / inner, in_ = l % inner;
        mo_cpx *base = data + o * inner * n + in_;
        for (int m = 0; m < n; ++m) buf[m] = base[(long)m * inner];
        mo_fft_line(&p, buf, buf + n, sign);
        for (int m = 0; m < n; ++m) base[(long)m * inner] = buf[m];
      }
      free(buf);
    }
  }
  mo_fplan_free(&p);
  return err ? MO_ALLOC_ERROR : MO_OK;
}

/* src/fft/fftw_engine.cc:130-143 (r2c, unnormalised, input preserved) */
int mo_fft_r2c(int d, const int *res, int nc, const double *in, double *out_) {
  mo_cpx *out = (mo_cpx *)out_;
  int h[3];
  mo_hermitian_sizes(d, res, h);
  const int n = res[d - 1], nh = h[d - 1];
  long rows = 1;
  for (int a = 0; a < d - 1; ++a) rows *= res[a];
  const long nlines = rows * nc;
  mo_fplan p;
  if (mo_fplan_init(&p, n)) return MO_ALLOC_ERROR;
  int err = 0;
  /* lines are paired: line a in the real part, line b in the imaginary part
   * of one complex transform, then split by Hermitian symmetry */
#pragma omp parallel num_threads(g_threads)
  {
    mo_cpx *buf = (mo_cpx *)malloc(sizeof(mo_cpx) * 2 * (size_t)n);
    if (!buf) {
#pragma omp atomic write
      err = 1;
    } else {
#pragma omp for schedule(static)
      for (long pr = 0; pr < (nlines + 1) / 2; ++pr) {
        const long la = 2 * pr, lb = 2 * pr + 1;
        const long ra = la / nc, ca = la % nc;
        const double *xa = in + ra * n * nc + ca;
        const int has_b = lb < nlines;
        const long rb = lb / nc, cb = lb % nc;
        const double *xb = has_b ? in + rb * n * nc + cb : NULL;
        for (int m = 0; m < n; ++m) {
          buf[m].re = xa[(long)m * nc];
          buf[m].im = has_b ? xb[(long)m * nc] : 0.0;
        }
        mo_fft_line(&p, buf, buf + n, -1);
        mo_cpx *ya = out + ra * nh * nc + ca;
        mo_cpx *yb = has_b ? out + rb * nh * nc + cb : NULL;
        for (int k = 0; k < nh; ++k) {
          mo_cpx z = buf[k], zc = buf[(n - k) % n];
          /* Xa = (Z[k] + conj Z[n-k]) / 2 ; Xb = (Z[k] - conj Z[n-k]) / 2i */
          mo_cpx A = {0.5 * (z.re + zc.re), 0.5 * (z.im - zc.im)};
          ya[(long)k * nc] = A;
          if (has_b) {
            mo_cpx B = {0.5 * (z.im + zc.im), -0.5 * (z.re - zc.re)};
            yb[(long)k * nc] = B;
          }
        }
      }
      free(buf);
    }
  }
  mo_fplan_free(&p);
  if (err) return MO_ALLOC_ERROR;
  for (int ax = d - 2; ax >= 0; --ax) {
    int e = c2c_axis(d, h, nc, out, ax, -1);
    if (e) return e;
  }
  return MO_OK;
}

/* src/fft/fftw_engine.cc:146-158 (c2r, unnormalised).  The input is not
 * modified here (the reference's FFTW plan overwrites its work field; the
 * callers never rely on that).  Axes 0..d-2 are inverted first, then each
 * last-axis line is completed by Hermitian symmetry; the imaginary parts
 * of the k=0 and (even n) k=n/2 entries are discarded, as in numpy's
 * irfftn. */
int mo_fft_c2r(int d, const int *res, int nc, const double *in_, double *out) {
  int h[3];
  mo_hermitian_sizes(d, res, h);
  const long kpts = mo_grid_size(d, h);
  mo_cpx *w = (mo_cpx *)malloc(sizeof(mo_cpx) * (size_t)(kpts * nc));
  if (!w) return MO_ALLOC_ERROR;
  memcpy(w, in_, sizeof(mo_cpx) * (size_t)(kpts * nc));
  for (int ax = 0; ax < d - 1; ++ax) {
    int e = c2c_axis(d, h, nc, w, ax, +1);
    if (e) { free(w); return e; }
  }
  const int n = res[d - 1], nh = h[d - 1];
  long rows = 1;
  for (int a = 0; a < d - 1; ++a) rows *= res[a];
  const long nlines = rows * nc;
  mo_fplan p;
  if (mo_fplan_init(&p, n)) { free(w); return MO_ALLOC_ERROR; }
  int err = 0;
#pragma omp parallel num_threads(g_threads)
  {
    mo_cpx *buf = (mo_cpx *)malloc(sizeof(mo_cpx) * 2 * (size_t)n);
    if (!buf) {
#pragma omp atomic write
      err = 1;
    } else {
#pragma omp for schedule(static)
      for (long pr = 0; pr < (nlines + 1) / 2; ++pr) {
        const long la = 2 * pr, lb = 2 * pr + 1;
        const long ra = la / nc, ca = la % nc;
        const int has_b = lb < nlines;
        const long rb = lb / nc, cb = lb % nc;
        const mo_cpx *Xa = w + ra * nh * nc + ca;
        const mo_cpx *Xb = has_b ? w + rb * nh * nc + cb : NULL;
        for (int k = 0; k < n; ++k) {
          mo_cpx a, b = {0.0, 0.0};
          int kk = k;
          int conj = 0;
          if (k >= nh) { kk = n - k; conj = 1; }
          a = Xa[(long)kk * nc];
          if (has_b) b = Xb[(long)kk * nc];
          if (kk == 0 || (n % 2 == 0 && kk == n / 2)) {
            a.im = 0.0;
            b.im = 0.0;
          }
          if (conj) { a.im = -a.im; b.im = -b.im; }
          /* Z = A + i B */
          buf[k].re = a.re - b.im;
          buf[k].im = a.im + b.re;
        }
        mo_fft_line(&p, buf, buf + n, +1);
        double *ya = out + ra * n * nc + ca;
        double *yb = has_b ? out + rb * n * nc + cb : NULL;
        for (int m = 0; m < n; ++m) {
          ya[(long)m * nc] = buf[m].re;
          if (has_b) yb[(long)m * nc] = buf[m].im;
        }
      }
      free(buf);
    }
  }
  mo_fplan_free(&p);
  free(w);
  return err ? MO_ALLOC_ERROR : MO_OK;
}
