/* Internal helpers of the CPU oracle (test infrastructure only). */
#ifndef MO_INTERNAL_H
#define MO_INTERNAL_H

#include <stddef.h>

typedef struct {
  double re, im;
} mo_cpx;

/* one 1D complex FFT length: factors and the table W_n^j = exp(-2*pi*i*j/n) */
typedef struct {
  int n;
  int nf;
  int f[64];
  mo_cpx *tw;
} mo_fplan;

int mo_fplan_init(mo_fplan *p, int n);
void mo_fplan_free(mo_fplan *p);
/* in-place transform of a contiguous line; scratch holds n entries.
 * sign = -1 forward, +1 backward (unnormalised). */
void mo_fft_line(const mo_fplan *p, mo_cpx *x, mo_cpx *scratch, int sign);

void mo_set_error(const char *fmt, ...);

/* tensor index helpers (column-major 2nd order, T4 map) */
#define T2(A, i, j, d) ((A)[(i) + (d) * (j)])
#define T4(A, i, j, k, l, d) \
  ((A)[((i) + (d) * (j)) + (d) * (d) * ((k) + (d) * (l))])

#endif
