/*
 * Host C layer under AddressSanitizer and UndefinedBehaviorSanitizer (built
 * and run by tests/test_host_sanitizers.py with gcc, CPU only).
 *
 * - pacg_host.c: twiddle tables, digit-reversal tables, the swap permutation
 *   and misc helpers, over many sizes;
 * - pacg_batch.c (included, so its static item code is reachable): batch
 *   creation, item building and the longest-processing-time dealing, with
 *   the HIP glue stubbed by host memory, checked for coverage and grouping.
 *
 * The variant table comes from the built library (pacg_batch_variants.h,
 * written by the test next to this file's object).
 */
#define _GNU_SOURCE
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../c-permutation-avoiding-convolution_amd/csrc/pacg_batch.c"
#include "pacg_fft.h"

#include "pacg_batch_variants.h"

/* ---- stubs of the HIP glue: "device" memory is host memory ---------------- */
static char g_err[512];
int pacg_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
void pacg_clear_error(void) { g_err[0] = 0; }
int pacg_dev_malloc(void** p, size_t bytes) {
  *p = malloc(bytes ? bytes : 1);
  return *p ? 0 : 2;
}
int pacg_dev_free(void* p) {
  free(p);
  return 0;
}
int pacg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  (void)stream;
  memcpy(dst, src, bytes);
  return 0;
}
int pacg_stream_sync(void* s) { return s ? 0 : 0; }
int pacg_stream_create(void** s) {
  *s = malloc(1);
  return 0;
}
int pacg_stream_destroy(void* s) {
  free(s);
  return 0;
}
int pacg_cu_count(int* cus) {
  *cus = 256;
  return 0;
}
int pacg_gen_twiddles(void* out, int prec, uint32_t N, uint32_t count, uint32_t step, void* stream) {
  (void)out, (void)prec, (void)N, (void)count, (void)step, (void)stream;
  return 0;
}
/* the makespan of the dealt items is the stand-in for a launch's time */
static double g_last_makespan;
int pacg_batch_launch(const pacg_batch_args_t* a, int blocks, void* stream) {
  (void)stream;
  double worst = 0.0;
  for (int w = 0; w < blocks; ++w) {
    double t = 0.0;
    for (uint32_t i = a->wg_first[w]; i < a->wg_first[w + 1]; ++i) t += item_cost(a->items[i].log2n, a->items[i].variant);
    if (t > worst) worst = t;
  }
  g_last_makespan = worst;
  return 0;
}
int pacg_event_create(void** ev) {
  *ev = malloc(sizeof(double));
  return 0;
}
int pacg_event_destroy(void* ev) {
  free(ev);
  return 0;
}
int pacg_event_record(void* ev, void* stream) {
  (void)stream;
  *(double*)ev = g_last_makespan;
  return 0;
}
int pacg_event_elapsed_ms(float* ms, void* a, void* b) {
  (void)a;
  *ms = (float)(*(double*)b * 1e-4);
  return 0;
}
int pacg_graph_begin(void* s) { return s ? 0 : 0; }
int pacg_graph_end(void* s, void** exec) {
  (void)s;
  *exec = malloc(1);
  return 0;
}
int pacg_graph_launch(void* e, void* s) { return (e && s) ? 0 : 0; }
int pacg_graph_destroy(void* e) {
  free(e);
  return 0;
}
const char* pacg_hip_error_string(int code) { return code ? "stub error" : "no error"; }

/* ---- checks ---------------------------------------------------------------- */
static int g_fail;
#define CHECK(c, ...)                            \
  do {                                           \
    if (!(c)) {                                  \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);              \
      fprintf(stderr, "\n");                     \
      ++g_fail;                                  \
    }                                            \
  } while (0)

static void check_batch(pacg_batch* b, const char* what) {
  const pacg_batch_item_t* it = b->d_items;
  const uint32_t* wf = b->d_wg_first;
  int* seen = calloc((size_t)b->count, sizeof(int));
  CHECK(b->n_wg >= 1 && b->n_wg <= b->blocks, "%s: %d workgroups", what, b->n_wg);
  CHECK(wf[0] == 0 && wf[b->n_wg] == (uint32_t)b->n_items, "%s: ranges", what);
  double total = 0.0, worst_item = 0.0, worst_wg = 0.0;
  for (int w = 0; w < b->n_wg; ++w) {
    CHECK(wf[w] < wf[w + 1], "%s: workgroup %d empty", what, w);
    double t = 0.0;
    for (uint32_t i = wf[w]; i < wf[w + 1]; ++i) {
      const double c = item_cost(it[i].log2n, it[i].variant);
      t += c;
      if (c > worst_item) worst_item = c;
      /* a shape appears in one run per workgroup */
      for (uint32_t j = i + 1; j < wf[w + 1]; ++j)
        if (it[j].log2n == it[i].log2n && it[j].variant == it[i].variant)
          CHECK(j == i + 1 || (it[j - 1].log2n == it[i].log2n && it[j - 1].variant == it[i].variant),
                "%s: shape 2^%d split in workgroup %d", what, it[i].log2n, w);
      const int per = per_item(it[i].log2n, it[i].variant);
      CHECK(it[i].count >= 1 && it[i].count <= per, "%s: item count %d of %d", what, it[i].count, per);
      CHECK(it[i].variant == b->variant[it[i].log2n], "%s: variant", what);
      for (uint32_t p = it[i].first; p < it[i].first + it[i].count; ++p) {
        CHECK(p < (uint32_t)b->count, "%s: position %u", what, p);
        if (p < (uint32_t)b->count) {
          seen[p]++;
          CHECK(b->log2n[b->order[p]] == it[i].log2n, "%s: pair length", what);
        }
      }
    }
    total += t;
    if (t > worst_wg) worst_wg = t;
  }
  for (int p = 0; p < b->count; ++p) CHECK(seen[p] == 1, "%s: position %d covered %d times", what, p, seen[p]);
  if (b->dealing == PACG_DEAL_LPT) {
    /* longest processing time first: within one item of the mean */
    CHECK(worst_wg <= total / b->n_wg + worst_item + 1e-9, "%s: makespan %.0f, mean %.0f, item %.0f", what,
          worst_wg, total / b->n_wg, worst_item);
  } else {
    /* slices of the built order: a workgroup's mixed-batch cost (plus a run
     * cost per slice) is within two items and a run cost of its share */
    double tm = 0.0, wm = 0.0, im = 0.0;
    for (int w = 0; w < b->n_wg; ++w) {
      double t = RUN_COST;
      for (uint32_t i = wf[w]; i < wf[w + 1]; ++i) {
        const double c = item_cost_mixed(it[i].log2n, it[i].variant);
        t += c;
        if (c > im) im = c;
      }
      tm += t;
      if (t > wm) wm = t;
    }
    CHECK(wm <= tm / b->n_wg + 2.0 * im + RUN_COST + 1e-9, "%s: slice makespan %.0f, mean %.0f, item %.0f", what,
          wm, tm / b->n_wg, im);
  }
  free(seen);
}

static void batches(void) {
  unsigned s = 12345u;
  const int counts[] = {1, 2, 7, 255, 256, 257, 1000, 4096, 9000};
  for (unsigned c = 0; c < sizeof counts / sizeof counts[0]; ++c) {
    const int n = counts[c];
    int* m = malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) {
      s = s * 1103515245u + 12345u;
      m[i] = 8 + (int)((s >> 16) % 7u);
    }
    if (c == 3)
      for (int i = 0; i < n; ++i) m[i] = 13;  /* one length only */
    pacg_batch* b = NULL;
    CHECK(pacg_batch_create(&b, n, m) == 0 && b, "create %d pairs: %s", n, g_err);
    if (b) {
      char what[64];
      snprintf(what, sizeof what, "%d pairs", n);
      check_batch(b, what);
      for (int len = LMIN; len <= LMAX; ++len)
        for (int v = 1; v < pacg_batch_num_variants(len); ++v) {
          CHECK(pacg_batch_set_variant(b, len, v) == 0, "set variant");
          check_batch(b, what);
          CHECK(pacg_batch_set_variant(b, len, 0) == 0, "set variant 0");
        }
      CHECK(pacg_batch_set_dealing(b, PACG_DEAL_SLICES) == 0 && pacg_batch_get_dealing(b) == PACG_DEAL_SLICES,
            "set dealing: %s", g_err);
      check_batch(b, what);
      CHECK(pacg_batch_set_dealing(b, PACG_DEAL_LPT) == 0, "set dealing back");
      CHECK(pacg_batch_set_dealing(b, 7) == PACG_ERR_ARG, "dealing 7 accepted");
      float fake = 0.0f;
      CHECK(pacg_batch_autotune(b, &fake, &fake, &fake, &fake, &fake, &fake, 2) >= 0, "autotune: %s", g_err);
      check_batch(b, what);
      CHECK(pacg_batch_capture(b, &fake, &fake, &fake, &fake, &fake, &fake) == 0, "capture");
      pacg_batch_destroy(b);
    }
    free(m);
  }
  /* rejected inputs */
  int bad[2] = {7, 10};
  pacg_batch* b = NULL;
  CHECK(pacg_batch_create(&b, 2, bad) == PACG_ERR_SIZE && !b, "length 2^7 accepted");
  bad[0] = 15;
  CHECK(pacg_batch_create(&b, 2, bad) == PACG_ERR_SIZE && !b, "length 2^15 accepted");
  const int big = 1 << 16;
  int* m = malloc(big * sizeof(int));
  for (int i = 0; i < big; ++i) m[i] = 14;  /* 2^30 points */
  CHECK(pacg_batch_create(&b, big, m) == PACG_ERR_SIZE && !b, "2^30 points accepted");
  free(m);
  CHECK(pacg_batch_create(&b, 0, bad) == PACG_ERR_ARG, "empty batch accepted");
}

static void host_tables(void) {
  for (int t = 1; t <= 16; ++t) {
    const int n = 1 << t;
    for (int r = 2; r <= 8; r *= 2) {
      if ((r == 4 && t % 2) || (r == 8 && t % 3)) continue;
      double **tre, **tim;
      int tt = r == 2 ? t : (r == 4 ? t / 2 : t / 3);
      if (r == 2) precompute_twiddles_r2(n, &tre, &tim);
      else if (r == 4) precompute_twiddles_r4(n, &tre, &tim);
      else precompute_twiddles_r8(n, &tre, &tim);
      free_twiddles(tre, tim, tt);
      int* rho = r == 2 ? precompute_index_reversal_permutation_r2(n)
                        : (r == 4 ? precompute_index_reversal_permutation_r4(n) : precompute_index_reversal_permutation_r8(n));
      int* hit = calloc((size_t)n, sizeof(int));
      for (int j = 0; j < n; ++j) {
        CHECK(rho[j] >= 0 && rho[j] < n, "rho out of range");
        if (rho[j] >= 0 && rho[j] < n) hit[rho[j]]++;
      }
      for (int j = 0; j < n; ++j) CHECK(hit[j] == 1, "rho not a bijection (n %d r %d)", n, r);
      double* re = calloc((size_t)n, sizeof(double));
      double* im = calloc((size_t)n, sizeof(double));
      double* flat = calloc((size_t)2 * n, sizeof(double));
      initialize_data_split(re, im, flat, n);
      apply_permutation_split(re, im, n, rho);
      apply_permutation_split(re, im, n, rho);
      CHECK(get_max_error(re, im, flat, n) == 0.0, "permutation twice is not the identity");
      free(re), free(im), free(flat), free(hit), free(rho);
    }
  }
  for (int t = 1; t <= 7; ++t) {
    int n = 1 << t;
    free(precompute_index_reversal_permutation_r2_2d(n, 2 * n));
    free(precompute_index_reversal_permutation_r2_3d(n, 2, 4 * n));
    if (t % 2 == 0) {
      free(precompute_index_reversal_permutation_r4_2d(n, 4 * n));
      free(precompute_index_reversal_permutation_r4_3d(n, 4, n));
    }
    double* re = calloc((size_t)n * n, sizeof(double));
    double* im = calloc((size_t)n * n, sizeof(double));
    in_place_square_array_transpose_split_2d(re, im, n);
    free(re), free(im);
  }
  int radix[5] = {4, 8, 8, 2, 4};
  for (long long p = 0; p < 2048; ++p) CHECK(pacg_rev_index_mixed(p, 2048, radix, 5) < 2048, "mixed reversal");
}

int main(void) {
  host_tables();
  batches();
  if (g_fail) {
    fprintf(stderr, "%d checks failed\n", g_fail);
    return 1;
  }
  printf("OK\n");
  return 0;
}
