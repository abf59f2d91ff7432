/*
 * Radix-schedule search for plans (host side, C11).
 *
 * The reference fixes one radix per call (convolution_3d.c:36-72 runs every
 * axis at radix r); a plan takes any mixed schedule of radix-2/4/8 stages per
 * axis, and which one is fastest depends on the compiled pass kernels, the
 * axis' stride and the precision.  pacg_plan_search tries, axis by axis, the
 * candidate schedules of that axis' length with the other axes fixed
 * (coordinate descent, two sweeps), each as a plan of its own with its tile
 * autotuner run, and keeps the fastest convolution; then, unless the caller
 * chose the axis of the fused pass, it tries each axis for it.
 *
 * Candidates of a length 2^t: every count (n8, n4, n2) of radix-8, -4 and -2
 * stages with 3 n8 + 2 n4 + n2 = t and at most one radix-2 stage, each in two
 * orders (radix descending from stage 0, and ascending).
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pacg.h"
#include "pacg_internal.h"

/* One candidate per row of out (max_stages ints each, the unused tail 0);
 * returns how many there are (up to max_cand are written), or -error. */
int pacg_schedule_candidates(int n, int* out, int max_cand, int max_stages) {
  int t = 0;
  while ((1LL << t) < n) ++t;
  if (n < 1 || (1LL << t) != n) return -pacg_fail(PACG_ERR_SIZE, "length %d is not a power of two", n);
  if (max_cand < 0 || max_stages < 1 || (max_cand > 0 && !out))
    return -pacg_fail(PACG_ERR_ARG, "pacg_schedule_candidates: bad argument");
  int count = 0;
  for (int n8 = t / 3; n8 >= 0; --n8)
    for (int n2 = 0; n2 <= 1; ++n2) {
      const int rest = t - 3 * n8 - n2;
      if (rest < 0 || rest % 2) continue;
      const int n4 = rest / 2, ns = n8 + n4 + n2;
      if (ns == 0 || ns > max_stages || ns > PACG_MAX_STAGES) continue;
      int desc[PACG_MAX_STAGES], k = 0;
      for (int i = 0; i < n8; ++i) desc[k++] = 8;
      for (int i = 0; i < n4; ++i) desc[k++] = 4;
      for (int i = 0; i < n2; ++i) desc[k++] = 2;
      for (int order = 0; order < 2; ++order) {
        int row[PACG_MAX_STAGES];
        for (int i = 0; i < ns; ++i) row[i] = order ? desc[ns - 1 - i] : desc[i];
        /* the ascending order of a single radix is the descending one */
        if (order && ns > 0 && row[0] == desc[0]) {
          int same = 1;
          for (int i = 0; i < ns; ++i) same &= row[i] == desc[i];
          if (same) continue;
        }
        if (count < max_cand) {
          int* dst = out + (size_t)count * (size_t)max_stages;
          memset(dst, 0, (size_t)max_stages * sizeof(int));
          memcpy(dst, row, (size_t)ns * sizeof(int));
        }
        ++count;
      }
    }
  return count;
}

/* xorshift64* in [-1, 1) */
static double rnd_pm1(uint64_t* s) {
  *s ^= *s >> 12;
  *s ^= *s << 25;
  *s ^= *s >> 27;
  return (double)((*s * 0x2545F4914F6CDD1DULL) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

typedef struct {
  void *xr, *xi, *yr, *yi; /* plan precision */
  double *kr, *ki;         /* FP64, device */
} search_bufs_t;

/* Milliseconds per convolution of the plan described by d, tiles autotuned;
 * a negative value when no plan can be made for d (no kernel, bad schedule). */
static double time_desc(const pacg_desc* d, const search_bufs_t* b, int reps) {
  pacg_plan* pl = NULL;
  if (pacg_plan_create(&pl, d) != PACG_OK) return -1.0;
  double ms = -1.0;
  void *e0 = NULL, *e1 = NULL;
  if (pacg_set_kernel(pl, b->kr, b->ki, 1) == PACG_OK && pacg_plan_autotune(pl, b->xr, b->xi, b->yr, b->yi, 2) >= 0 &&
      pacg_event_create(&e0) == 0 && pacg_event_create(&e1) == 0) {
    void* st = pacg_plan_stream(pl);
    int ok = pacg_convolve(pl, b->xr, b->xi, b->yr, b->yi) == PACG_OK;
    if (ok && pacg_event_record(e0, st) == 0) {
      for (int r = 0; r < reps && ok; ++r) ok = pacg_convolve(pl, b->xr, b->xi, b->yr, b->yi) == PACG_OK;
      float t = 0.0f;
      if (ok && pacg_event_record(e1, st) == 0 && pacg_stream_sync(st) == 0 && pacg_event_elapsed_ms(&t, e0, e1) == 0)
        ms = (double)t / reps;
    }
  }
  if (e0) pacg_event_destroy(e0);
  if (e1) pacg_event_destroy(e1);
  pacg_plan_destroy(pl);
  return ms;
}

int pacg_plan_search(pacg_desc* d, int reps) {
  if (!d || d->nd < 1 || d->nd > 3 || reps < 1) return -pacg_fail(PACG_ERR_ARG, "pacg_plan_search: bad argument");
  long long n = 1;
  for (int a = 0; a < d->nd; ++a) {
    if (d->dims[a] < 1) return -pacg_fail(PACG_ERR_SIZE, "axis %d length %d", a, d->dims[a]);
    n *= d->dims[a];
  }
  if (n >= (1LL << 31)) return -pacg_fail(PACG_ERR_SIZE, "total size %lld >= 2^31", n);
  pacg_desc cur = *d;
  const size_t es = cur.precision == PACG_PREC_F32 ? 4u : 8u;
  search_bufs_t b;
  memset(&b, 0, sizeof b);
  int rc = PACG_OK, changed = 0;
  double* host = (double*)malloc((size_t)n * sizeof(double));
  if (!host) {
    rc = pacg_fail(PACG_ERR_NOMEM, "out of host memory");
    goto out;
  }
  if (pacg_dev_malloc(&b.xr, (size_t)n * es) || pacg_dev_malloc(&b.xi, (size_t)n * es) ||
      pacg_dev_malloc(&b.yr, (size_t)n * es) || pacg_dev_malloc(&b.yi, (size_t)n * es) ||
      pacg_dev_malloc((void**)&b.kr, (size_t)n * 8u) || pacg_dev_malloc((void**)&b.ki, (size_t)n * 8u)) {
    rc = pacg_fail(PACG_ERR_NOMEM, "device allocation for the schedule search failed");
    goto out;
  }
  {
    /* seeded random x (in the plan precision) and k (FP64), uploaded once */
    uint64_t s = 0x9E3779B97F4A7C15ULL;
    void* dst[4] = {b.xr, b.xi, b.kr, b.ki};
    for (int q = 0; q < 4; ++q) {
      const int f64 = q >= 2 || es == 8u;
      for (long long i = 0; i < n; ++i) {
        const double v = rnd_pm1(&s);
        if (f64)
          host[i] = v;
        else
          ((float*)host)[i] = (float)v;
      }
      if (pacg_memcpy_h2d(dst[q], host, (size_t)n * (f64 ? 8u : 4u), NULL) || pacg_stream_sync(NULL)) {
        rc = pacg_fail(PACG_ERR_HIP, "upload for the schedule search failed");
        goto out;
      }
    }
  }
  /* the starting point: the caller's schedules (defaults where none is given) */
  double best = time_desc(&cur, &b, reps);
  if (best < 0) {
    rc = pacg_fail(PACG_ERR_NOKERNEL, "pacg_plan_search: the starting schedules make no plan");
    goto out;
  }
  enum { MAXC = 32 };
  int cand[MAXC * PACG_MAX_STAGES];
  for (int sweep = 0; sweep < 2; ++sweep)
    for (int a = 0; a < cur.nd; ++a) {
      if (cur.nstages[a] < 0 || cur.dims[a] == 1) continue; /* a batch axis, or nothing to transform */
      int nc = pacg_schedule_candidates(cur.dims[a], cand, MAXC, PACG_MAX_STAGES);
      if (nc < 0) {
        rc = -nc;
        goto out;
      }
      if (nc > MAXC) nc = MAXC;
      for (int c = 0; c < nc; ++c) {
        pacg_desc trial = cur;
        int ns = 0;
        while (ns < PACG_MAX_STAGES && cand[c * PACG_MAX_STAGES + ns]) ++ns;
        trial.nstages[a] = ns;
        for (int q = 0; q < ns; ++q) trial.radix[a][q] = cand[c * PACG_MAX_STAGES + q];
        if (cur.nstages[a] == ns && memcmp(trial.radix[a], cur.radix[a], (size_t)ns * sizeof(int)) == 0) continue;
        const double ms = time_desc(&trial, &b, reps);
        /* a new schedule must beat the current one by more than 1 % (no flips on noise) */
        if (ms > 0 && ms < 0.99 * best) {
          best = ms;
          cur = trial;
          changed |= 1 << a;
        }
      }
    }
  /* the axis of the fused pass, unless the caller chose one: each transformed
   * axis in turn with the schedules found (-1, the plan's default rule, stays
   * unless an explicit axis is more than 1 % faster) */
  if (d->fused_axis < 0)
    for (int a = 0; a < cur.nd; ++a) {
      if (cur.nstages[a] < 0 || cur.dims[a] == 1) continue;
      pacg_desc trial = cur;
      trial.fused_axis = a;
      const double ms = time_desc(&trial, &b, reps);
      if (ms > 0 && ms < 0.99 * best) {
        best = ms;
        cur = trial;
      }
    }
  for (int a = 0; a < cur.nd; ++a) {
    d->nstages[a] = cur.nstages[a];
    memcpy(d->radix[a], cur.radix[a], sizeof cur.radix[a]);
  }
  d->fused_axis = cur.fused_axis;
  pacg_clear_error();
out:
  free(host);
  pacg_dev_free(b.xr);
  pacg_dev_free(b.xi);
  pacg_dev_free(b.yr);
  pacg_dev_free(b.yi);
  pacg_dev_free(b.kr);
  pacg_dev_free(b.ki);
  return rc ? -rc : __builtin_popcount((unsigned)changed);
}
