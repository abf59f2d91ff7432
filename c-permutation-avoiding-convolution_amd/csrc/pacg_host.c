/*
 * Host helpers with the reference's names and semantics: twiddle tables in
 * the reference's jagged layout, base-r digit reversal and its 2D/3D
 * composition, the swap permutation, and misc.c's utilities.  Integer
 * bookkeeping is bit-exact with the reference by construction (same closed
 * forms); it runs on the host because callers hand these tables back to
 * host code.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pacg.h"
#include "pacg_host.h"

long long pacg_rev_index_mixed(long long p, long long n, const int* radix, int t) {
  long long out = 0, l = 1;
  for (int q = 0; q < t; ++q) {
    long long d = (p / l) % radix[q];
    long long k = l * radix[q];
    out += d * (n / k);
    l = k;
  }
  return out;
}

/* number of base-r digits of n (the reference's "while (temp > 1) temp >>= b")
 * -- twiddle_factors.c:40-45 */
static int digits_r(int n, int r) {
  int bits = r == 2 ? 1 : r == 4 ? 2 : 3, t = 0;
  for (int m = n; m > 1; m >>= bits) ++t;
  return t;
}

/* Jagged FP64 tables: level q holds w_k^{(m+1) j}, m < r-1, j < ell = r^q,
 * k = r^{q+1} at index m*ell + j (twiddle_factors.c:17-131). */
static void twiddles_r(int n, int r, double*** tre, double*** tim) {
  int t = digits_r(n, r);
  double** wr = (double**)malloc((size_t)(t > 0 ? t : 1) * sizeof(double*));
  double** wi = (double**)malloc((size_t)(t > 0 ? t : 1) * sizeof(double*));
  long long ell = 1;
  for (int q = 0; q < t; ++q, ell *= r) {
    long long k = ell * r;
    wr[q] = (double*)malloc((size_t)((r - 1) * ell) * sizeof(double));
    wi[q] = (double*)malloc((size_t)((r - 1) * ell) * sizeof(double));
    for (int m = 0; m < r - 1; ++m)
      for (long long j = 0; j < ell; ++j) {
        double angle = -2.0 * M_PI * (double)((m + 1) * j) / (double)k;
        wr[q][m * ell + j] = cos(angle);
        wi[q][m * ell + j] = sin(angle);
      }
  }
  *tre = wr;
  *tim = wi;
}

void precompute_twiddles_r2(int n, double*** tre, double*** tim) { twiddles_r(n, 2, tre, tim); }
void precompute_twiddles_r4(int n, double*** tre, double*** tim) { twiddles_r(n, 4, tre, tim); }
void precompute_twiddles_r8(int n, double*** tre, double*** tim) { twiddles_r(n, 8, tre, tim); }

void free_twiddles(double** tre, double** tim, int t) {
  if (tre)
    for (int q = 0; q < t; ++q) free(tre[q]);
  if (tim)
    for (int q = 0; q < t; ++q) free(tim[q]);
  free(tre);
  free(tim);
}

/* base-r reversal over t digits (index_reversals.c:13-20,30-37,52-59) */
static int reverse_bits_group(int x, int t, int bits) {
  int y = 0, mask = (1 << bits) - 1;
  for (int i = 0; i < t; ++i) {
    y = (y << bits) | (x & mask);
    x >>= bits;
  }
  return y;
}
int reverse_index_r2(int x, int t) { return reverse_bits_group(x, t, 1); }
int reverse_index_r4(int x, int t) { return reverse_bits_group(x, t, 2); }
int reverse_index_r8(int x, int t) { return reverse_bits_group(x, t, 3); }

static int* rho_nd(int nd, const int* dims, int r) {
  int bits = r == 2 ? 1 : r == 4 ? 2 : 3;
  long long n = 1;
  int t[3];
  for (int a = 0; a < nd; ++a) {
    n *= dims[a];
    t[a] = digits_r(dims[a], r);
  }
  int* rho = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  if (!rho) return NULL;
  for (long long idx = 0; idx < n; ++idx) {
    long long rem = idx, acc = 0, mul = 1;
    for (int a = nd - 1; a >= 0; --a) {
      int i = (int)(rem % dims[a]);
      rem /= dims[a];
      acc += (long long)reverse_bits_group(i, t[a], bits) * mul;
      mul *= dims[a];
    }
    rho[idx] = (int)acc;
  }
  return rho;
}

/* index_reversals.c:22-28,39-50,61-72; _2d.c:14-62; _3d.c:14-72 */
int* precompute_index_reversal_permutation_r2(int n) { return rho_nd(1, &n, 2); }
int* precompute_index_reversal_permutation_r4(int n) { return rho_nd(1, &n, 4); }
int* precompute_index_reversal_permutation_r8(int n) { return rho_nd(1, &n, 8); }
int* precompute_index_reversal_permutation_r2_2d(int n1, int n2) {
  int d[2] = {n1, n2};
  return rho_nd(2, d, 2);
}
int* precompute_index_reversal_permutation_r4_2d(int n1, int n2) {
  int d[2] = {n1, n2};
  return rho_nd(2, d, 4);
}
int* precompute_index_reversal_permutation_r2_3d(int n1, int n2, int n3) {
  int d[3] = {n1, n2, n3};
  return rho_nd(3, d, 2);
}
int* precompute_index_reversal_permutation_r4_3d(int n1, int n2, int n3) {
  int d[3] = {n1, n2, n3};
  return rho_nd(3, d, 4);
}

/* swap j <-> rho[j] where j < rho[j] (index_reversals.c:74-94) */
void apply_permutation(double complex* x, int n, const int* rho) {
  if (!x || !rho) return;
  for (int j = 0; j < n; ++j) {
    int r = rho[j];
    if (j < r) {
      double complex tmp = x[j];
      x[j] = x[r];
      x[r] = tmp;
    }
  }
}
void apply_permutation_split(double* restrict re, double* restrict im, int n, const int* rho) {
  if (!re || !im || !rho) return;
  for (int j = 0; j < n; ++j) {
    int r = rho[j];
    if (j < r) {
      double a = re[j], b = im[j];
      re[j] = re[r];
      im[j] = im[r];
      re[r] = a;
      im[r] = b;
    }
  }
}

/* misc.c:16-79 */
double time_diff(struct timespec start, struct timespec end) {
  return (double)(end.tv_sec - start.tv_sec) + (double)(end.tv_nsec - start.tv_nsec) / 1e9;
}
double get_max_error(double* re, double* im, double* flat, int n) {
  double mx = 0.0;
  for (int i = 0; i < n; ++i) {
    double e = hypot(re[i] - flat[2 * i], im[i] - flat[2 * i + 1]);
    if (e > mx) mx = e;
  }
  return mx;
}
void initialize_data(double complex* x, double* flat, int n) {
  for (int i = 0; i < n; ++i) {
    double a = rand() / (double)RAND_MAX;
    double b = rand() / (double)RAND_MAX;
    x[i] = a + I * b;
    flat[2 * i] = a;
    flat[2 * i + 1] = b;
  }
}
void initialize_data_split(double* re, double* im, double* flat, int n) {
  for (int i = 0; i < n; ++i) {
    re[i] = rand() / (double)RAND_MAX;
    im[i] = rand() / (double)RAND_MAX;
    flat[2 * i] = re[i];
    flat[2 * i + 1] = im[i];
  }
}
void initialize_filter(double* g_re, double* g_im, int n) {
  for (int i = 0; i < n; ++i) {
    g_re[i] = rand() / (double)RAND_MAX;
    g_im[i] = rand() / (double)RAND_MAX;
  }
}
bool supports_radix_4(int t2) { return (t2 % 2) == 0; }
bool supports_radix_8(int t2) { return (t2 % 3) == 0; }
void in_place_square_array_transpose_split_2d(double* re, double* im, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      long long a = (long long)i * n + j, b = (long long)j * n + i;
      double tr = re[a], ti = im[a];
      re[a] = re[b];
      im[a] = im[b];
      re[b] = tr;
      im[b] = ti;
    }
}
