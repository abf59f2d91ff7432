/*
 * bench_driver.c -- verify / bench programs over the drop-in fft.h entry points.
 *
 * One source for the eleven driver programs of the reference, built once per
 * program with
 *     -DPACG_DIM=1|2|3  -DPACG_KIND=PACG_CONV|PACG_FWD|PACG_BWD  [-DPACG_ANISO]
 * into bench{1,2,3}d_{convolution,forward_fft,backward_fft} and
 * bench{2,3}d_convolution_anisotropic.  Every compute call goes through the
 * reference's own names (fft.h), so the programs exercise libpacg.so exactly as a
 * reference caller linked against it would.
 *
 * Command line, defaults and output lines follow the reference programs
 * (bench1d_convolution.c:59-85, :264-268; bench1d_forward_fft.c:161-169;
 * bench2d_forward_fft.c:154-162; bench3d_convolution.c:73-88, :246-250, :314 of
 * the anisotropic one; and the matching lines of the others), so the reference's
 * bench*.sh sweeps run unchanged:
 *     prog verify|bench [-t power] [-r timing_runs]
 * -t sets log2 of the leading axis (default 24 in 1D, 12 in 2D, 8 in 3D), the other
 * axes are equal to it, or 2^2 in the anisotropic programs; -r sets the timing runs
 * (default 10).  Radix 4 runs when every axis has an even log2 and, in 1D, radix 8
 * when it is a multiple of 3.
 *
 * Two departures, both forced by what this library does not contain:
 *   - verify compares against the CPU FFT below instead of FFTW, and its lines end
 *     in "vs CPU FFT: <max abs error>";
 *   - bench prints no FFTW line.
 * bench times each call with CLOCK_MONOTONIC around it, as the reference does, so
 * host<->device copies are inside the figure (the drop-in takes host arrays).
 */
#define _POSIX_C_SOURCE 199309L
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "fft.h"
#include "pacg.h"

#define PACG_CONV 0
#define PACG_FWD 1
#define PACG_BWD 2

#ifndef PACG_DIM
#error "define PACG_DIM (1, 2 or 3)"
#endif
#ifndef PACG_KIND
#error "define PACG_KIND (PACG_CONV, PACG_FWD or PACG_BWD)"
#endif
#ifdef PACG_ANISO
#define ANISO 1
#else
#define ANISO 0
#endif

/* ---- problem and per-radix tables ------------------------------------------------ */

typedef struct {
  int t[3];    /* log2 of each axis */
  int n[3];    /* axis lengths */
  int npts;    /* product */
} shape_t;

typedef struct {
  int r;
  int levels[3];        /* table depth per axis, what free_twiddles needs */
  int* rho;             /* digit reversal of the whole array */
  double** twr[3];
  double** twi[3];
} tables_t;

static void tables_make(tables_t* tb, const shape_t* sh, int r) {
  memset(tb, 0, sizeof(*tb));
  tb->r = r;
  int bits = r == 2 ? 1 : r == 4 ? 2 : 3;
  for (int a = 0; a < PACG_DIM; ++a) {
    tb->levels[a] = sh->t[a] / bits;
    if (r == 2) precompute_twiddles_r2(sh->n[a], &tb->twr[a], &tb->twi[a]);
    else if (r == 4) precompute_twiddles_r4(sh->n[a], &tb->twr[a], &tb->twi[a]);
    else precompute_twiddles_r8(sh->n[a], &tb->twr[a], &tb->twi[a]);
  }
#if PACG_DIM == 1
  tb->rho = r == 2 ? precompute_index_reversal_permutation_r2(sh->n[0])
          : r == 4 ? precompute_index_reversal_permutation_r4(sh->n[0])
                   : precompute_index_reversal_permutation_r8(sh->n[0]);
#elif PACG_DIM == 2
  tb->rho = r == 2 ? precompute_index_reversal_permutation_r2_2d(sh->n[0], sh->n[1])
                   : precompute_index_reversal_permutation_r4_2d(sh->n[0], sh->n[1]);
#else
  tb->rho = r == 2 ? precompute_index_reversal_permutation_r2_3d(sh->n[0], sh->n[1], sh->n[2])
                   : precompute_index_reversal_permutation_r4_3d(sh->n[0], sh->n[1], sh->n[2]);
#endif
}

static void tables_free(tables_t* tb) {
  for (int a = 0; a < PACG_DIM; ++a) free_twiddles(tb->twr[a], tb->twi[a], tb->levels[a]);
  free(tb->rho);
}

/* ---- the calls under test, by name of the reference entry point ------------------ */

static void check_call(const char* what) {
  const char* e = pacg_last_error();
  if (e && e[0]) {
    fprintf(stderr, "%s failed: %s\n", what, e);
    exit(1);
  }
}

#if PACG_DIM == 1
#define TW(tb) (tb)->twr[0], (tb)->twi[0]
#define DIMS(sh) (sh)->n[0]
#elif PACG_DIM == 2
#define TW(tb) (tb)->twr[0], (tb)->twi[0], (tb)->twr[1], (tb)->twi[1]
#define DIMS(sh) (sh)->n[0], (sh)->n[1]
#else
#define TW(tb) (tb)->twr[0], (tb)->twi[0], (tb)->twr[1], (tb)->twi[1], (tb)->twr[2], (tb)->twi[2]
#define DIMS(sh) (sh)->n[0], (sh)->n[1], (sh)->n[2]
#endif

#if PACG_DIM == 1
#define CALL(name, ...) name(__VA_ARGS__)
#elif PACG_DIM == 2
#define CALL(name, ...) name##_2d(__VA_ARGS__)
#else
#define CALL(name, ...) name##_3d(__VA_ARGS__)
#endif

enum { STEP_A, STEP_B, STEP_C };  /* conv: FFT-based, PA; fft kinds: ordered, permutation, unordered */

static void run_step(int step, const shape_t* sh, const tables_t* tb, double* re, double* im, double* hr,
                     double* hi) {
  int r = tb->r;
#if PACG_KIND == PACG_CONV
  (void)sh;
  if (step == STEP_A) {
    CALL(fft_based_convolution, re, im, DIMS(sh), r, TW(tb), tb->rho, hr, hi);
    check_call("FFT-based convolution");
  } else {
    CALL(permutation_avoiding_convolution, re, im, DIMS(sh), r, TW(tb), hr, hi);
    check_call("permutation-avoiding convolution");
  }
#else
  (void)hr;
  (void)hi;
  if (step == STEP_A) {
#if PACG_KIND == PACG_FWD
    CALL(fft, re, im, DIMS(sh), r, TW(tb), tb->rho);
#else
    CALL(ifft, re, im, DIMS(sh), r, TW(tb), tb->rho);
#endif
    check_call("ordered transform");
  } else if (step == STEP_B) {
    apply_permutation_split(re, im, sh->npts, tb->rho);
  } else {
#if PACG_KIND == PACG_FWD
    CALL(ufft, re, im, DIMS(sh), r, TW(tb));
#else
    CALL(uifft, re, im, DIMS(sh), r, TW(tb));
#endif
    check_call("unordered transform");
  }
#endif
}

/* ---- CPU comparator: radix-2 FFT along every axis of an interleaved array -------- */

static const double TWO_PI = 6.283185307179586476925286766559;

static void cpu_fft_line(double complex* a, int m, int sign) {
  for (int i = 1, j = 0; i < m; ++i) {
    int bit = m >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double complex s = a[i];
      a[i] = a[j];
      a[j] = s;
    }
  }
  for (int len = 2; len <= m; len <<= 1) {
    int half = len >> 1;
    for (int k = 0; k < half; ++k) {
      double ang = sign * TWO_PI * (double)k / (double)len;
      double complex w = cos(ang) + I * sin(ang);
      for (int i = k; i < m; i += len) {
        double complex u = a[i], v = a[i + half] * w;
        a[i] = u + v;
        a[i + half] = u - v;
      }
    }
  }
}

/* sign -1: sum x_j e^{-2 pi i jk/n} (the forward transform); +1: unscaled inverse */
static void cpu_fft(double* flat, const shape_t* sh, int sign) {
  double complex* a = (double complex*)flat;
  int maxn = 1;
  for (int ax = 0; ax < PACG_DIM; ++ax) maxn = sh->n[ax] > maxn ? sh->n[ax] : maxn;
  double complex* line = malloc((size_t)maxn * sizeof(double complex));
  long stride = 1;
  for (int ax = PACG_DIM - 1; ax >= 0; --ax) {
    int m = sh->n[ax];
    long outer = (long)sh->npts / ((long)m * stride);
    for (long o = 0; o < outer; ++o)
      for (long s = 0; s < stride; ++s) {
        double complex* base = a + o * m * stride + s;
        if (stride == 1) {
          cpu_fft_line(base, m, sign);
          continue;
        }
        for (int j = 0; j < m; ++j) line[j] = base[j * stride];
        cpu_fft_line(line, m, sign);
        for (int j = 0; j < m; ++j) base[j * stride] = line[j];
      }
    stride *= m;
  }
  free(line);
}

#if PACG_KIND == PACG_CONV
/* flat <- (1/n) F^{-1}[(F flat) . (F h)] */
static void cpu_convolution(double* flat, const double* hr, const double* hi, const shape_t* sh) {
  size_t n = (size_t)sh->npts;
  double* hf = malloc(2 * n * sizeof(double));
  for (size_t i = 0; i < n; ++i) {
    hf[2 * i] = hr[i];
    hf[2 * i + 1] = hi[i];
  }
  cpu_fft(flat, sh, -1);
  cpu_fft(hf, sh, -1);
  double complex* x = (double complex*)flat;
  double complex* h = (double complex*)hf;
  for (size_t i = 0; i < n; ++i) x[i] *= h[i];
  cpu_fft(flat, sh, +1);
  for (size_t i = 0; i < 2 * n; ++i) flat[i] /= (double)n;
  free(hf);
}
#endif

/* ---- output lines ----------------------------------------------------------------- */

#if PACG_KIND == PACG_CONV
#if PACG_DIM == 3
#define WIDTH 49
static const char* const STEP_NAME[2] = {"FFT", "permutation-avoiding"};
#else
#define WIDTH 46
static const char* const STEP_NAME[2] = {"FFT-based", "permutation-avoiding"};
#endif
#define NSTEPS 2
#else
#if PACG_DIM == 1 || (PACG_DIM == 2 && PACG_KIND == PACG_BWD)
#define WIDTH 38
#else
#define WIDTH 40
#endif
static const char* const STEP_NAME[3] = {"FFT", "permutation", "unordered FFT"};
#define NSTEPS 3
#endif

static void print_time(int r, int step, double sum, double sum_sq, int runs) {
  const char* name = STEP_NAME[step];
#if PACG_KIND == PACG_CONV && PACG_DIM == 3 && ANISO
  if (r == 4 && step == STEP_B) name = "permutation-avoidning"; /* bench3d_convolution_anisotropic.c:314 */
#endif
  char label[96];
  snprintf(label, sizeof label, "Average radix-%d %s time:", r, name);
  double mu = sum / runs;
  double var = sum_sq / runs - mu * mu;
  printf("%-*s%.6f +/- %.6f seconds\n", WIDTH, label, mu, var > 0.0 ? sqrt(var) : 0.0);
}

/* ---- verify and bench ------------------------------------------------------------- */

typedef struct {
  double *re, *im, *flat, *h_re, *h_im, *hc_re, *hc_im;
} buffers_t;

static void restore_filter(const buffers_t* b, int n) {
  memcpy(b->hc_re, b->h_re, (size_t)n * sizeof(double));
  memcpy(b->hc_im, b->h_im, (size_t)n * sizeof(double));
}

static void verify(const shape_t* sh, const tables_t* tb, const buffers_t* b) {
  int n = sh->npts;
#if PACG_KIND == PACG_CONV
  static const char* const who[2] = {"FFT-based", "Permutation-avoiding"};
  for (int step = STEP_A; step <= STEP_B; ++step) {
    initialize_data_split(b->re, b->im, b->flat, n);
    cpu_convolution(b->flat, b->h_re, b->h_im, sh);
    restore_filter(b, n);
    run_step(step, sh, tb, b->re, b->im, b->hc_re, b->hc_im);
    printf("Max error %s radix-%d vs CPU FFT: %.3e\n", who[step], tb->r, get_max_error(b->re, b->im, b->flat, n));
  }
#else
  initialize_data_split(b->re, b->im, b->flat, n);
#if PACG_KIND == PACG_FWD
  cpu_fft(b->flat, sh, -1);
#else
  cpu_fft(b->flat, sh, +1);
  for (long i = 0; i < 2L * n; ++i) b->flat[i] /= (double)n;
#endif
  run_step(STEP_A, sh, tb, b->re, b->im, NULL, NULL);
  printf("Max error FFT radix-%d vs CPU FFT: %.3e\n", tb->r, get_max_error(b->re, b->im, b->flat, n));
#endif
}

static void bench(const shape_t* sh, const tables_t* tb, const buffers_t* b, int runs) {
  int n = sh->npts;
  double sum[NSTEPS] = {0}, sum_sq[NSTEPS] = {0};
  struct timespec t0, t1;
  for (int it = -1; it < runs; ++it) { /* it = -1: the untimed warm-up */
    initialize_data_split(b->re, b->im, b->flat, n);
    for (int step = 0; step < NSTEPS; ++step) {
      if (PACG_KIND == PACG_CONV) restore_filter(b, n);
      clock_gettime(CLOCK_MONOTONIC, &t0);
      run_step(step, sh, tb, b->re, b->im, b->hc_re, b->hc_im);
      clock_gettime(CLOCK_MONOTONIC, &t1);
      if (it >= 0) {
        double dt = time_diff(t0, t1);
        sum[step] += dt;
        sum_sq[step] += dt * dt;
      }
    }
  }
  for (int step = 0; step < NSTEPS; ++step) print_time(tb->r, step, sum[step], sum_sq[step], runs);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "Usage: %s [verify|bench] [-t power] [-r timing_runs]\n", argv[0]);
    return 1;
  }
  int do_verify = strcmp(argv[1], "verify") == 0;
  int do_bench = strcmp(argv[1], "bench") == 0;
  if (!do_verify && !do_bench) {
    fprintf(stderr, "Unknown mode: %s\n", argv[1]);
    return 1;
  }
  int t = PACG_DIM == 1 ? 24 : PACG_DIM == 2 ? 12 : 8;
  int runs = 10;
  for (int i = 2; i < argc; ++i) {
    if (strcmp(argv[i], "-t") == 0 && i + 1 < argc) {
      t = atoi(argv[++i]);
    } else if (strcmp(argv[i], "-r") == 0 && i + 1 < argc) {
      runs = atoi(argv[++i]);
    } else {
      fprintf(stderr, "Unknown or malformed option: %s\n", argv[i]);
      return 1;
    }
  }

  shape_t sh;
  long npts = 1;
  for (int a = 0; a < PACG_DIM; ++a) {
    sh.t[a] = a == 0 ? t : (ANISO ? 2 : t);
    if (sh.t[a] < 1 || sh.t[a] > 30) {
      fprintf(stderr, "power out of range: %d\n", sh.t[a]);
      return 1;
    }
    sh.n[a] = 1 << sh.t[a];
    npts *= sh.n[a];
  }
  if (npts > (1L << 30) || runs < 1) {
    fprintf(stderr, "unsupported size or run count\n");
    return 1;
  }
  sh.npts = (int)npts;

  size_t bytes = (size_t)npts * sizeof(double);
  buffers_t b = {malloc(bytes), malloc(bytes), malloc(2 * bytes), malloc(bytes),
                 malloc(bytes), malloc(bytes), malloc(bytes)};
  if (!b.re || !b.im || !b.flat || !b.h_re || !b.h_im || !b.hc_re || !b.hc_im) {
    fprintf(stderr, "out of host memory\n");
    return 1;
  }
  initialize_filter(b.h_re, b.h_im, sh.npts);

  int radices[3], nr = 0;
  int even = 1;
  for (int a = 0; a < PACG_DIM; ++a) even &= supports_radix_4(sh.t[a]);
  radices[nr++] = 2;
  if (even) radices[nr++] = 4;
  if (PACG_DIM == 1 && supports_radix_8(sh.t[0])) radices[nr++] = 8;

  for (int k = 0; k < nr; ++k) {
    tables_t tb;
    tables_make(&tb, &sh, radices[k]);
    if (do_verify) verify(&sh, &tb, &b);
    else bench(&sh, &tb, &b, runs);
    fflush(stdout);
    tables_free(&tb);
  }

  pacg_dropin_release();
  free(b.re);
  free(b.im);
  free(b.flat);
  free(b.h_re);
  free(b.h_im);
  free(b.hc_re);
  free(b.hc_im);
  return 0;
}
