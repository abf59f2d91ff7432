/*
 * Batched uneven-length convolutions, host side: packing, work items, their
 * assignment to workgroups, the per-length schedule autotuner and HIP-graph
 * capture of the single persistent launch (kernel: pacg_batch.hip).
 *
 * Pairs are packed back to back in input order (pair i at offset
 * sum_{j<i} 2^m_j of each of x.re, x.im, k.re, k.im, y.re, y.im).  The
 * kernel sees them sorted longest first and cut into items of 1024 / TPL
 * pairs of one length n (TPL threads per pair).  Items are dealt to the
 * workgroups longest processing time first (each to the least loaded one),
 * then each workgroup's items are ordered by shape, so that it meets runs
 * of one shape and can load the next item of a run ahead; or, where the
 * autotuner finds it faster, in contiguous slices of the shape-ordered list
 * (one or two shapes per workgroup, fewer run starts).
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>

#include "pacg.h"
#include "pacg_batch.h"
#include "pacg_host.h"
#include "pacg_internal.h"

#define LMIN 8
#define LMAX 14
#define BATCH_THREADS 1024
#define ITEM_POINTS (1 << 14)  /* points of the largest item (1024 threads x 16) */

struct pacg_batch {
  int count;
  int* log2n;           /* input order */
  long long* offset;    /* input order */
  long long total;
  int* order;           /* sorted position -> input index */
  int variant[LMAX + 1];
  int dealing;          /* PACG_DEAL_LPT or PACG_DEAL_SLICES */
  int blocks;           /* workgroups of the persistent launch */
  void* stream;
  int own_stream;
  void* graph;
  /* device */
  uint32_t* d_conv_off;
  pacg_batch_item_t* d_items;
  uint32_t* d_wg_first;  /* blocks + 1 entries */
  int n_items;
  int n_wg;              /* workgroups of the current item lists */
  int cap_items;
  void* d_tw;
};

static const pacg_batch_variant_t* variant_row(int m, int v) {
  for (int i = 0; i < pacg_num_batch_variants; ++i)
    if (pacg_batch_variants[i].log2n == m && pacg_batch_variants[i].variant == v) return &pacg_batch_variants[i];
  return NULL;
}

int pacg_batch_schedule(int log2n, int variant, int* radix, int maxn) {
  const pacg_batch_variant_t* r = variant_row(log2n, variant);
  if (!r) return -pacg_fail(PACG_ERR_ARG, "no batched variant %d for length 2^%d", variant, log2n);
  for (int q = 0; q < r->nstages && q < maxn; ++q) radix[q] = r->radix[q];
  return r->nstages;
}

int pacg_batch_num_variants(int log2n) {
  int c = 0;
  for (int i = 0; i < pacg_num_batch_variants; ++i) c += pacg_batch_variants[i].log2n == log2n;
  return c;
}

/* convolutions per item: the 1024 threads' worth of TPL-thread groups */
static int per_item(int m, int v) {
  const pacg_batch_variant_t* r = variant_row(m, v);
  if (!r || r->tpl < 1) return 1;
  int g = BATCH_THREADS / r->tpl;
  return g < 1 ? 1 : g;
}

/* Cost of an item: every thread of the workgroup runs the whole item (a
 * ragged item computes its missing pairs too and stores nothing), so it is the
 * variant's measured time per item, whatever the item's pair count. */
static double item_cost(int m, int v) {
  const pacg_batch_variant_t* r = variant_row(m, v);
  return r ? (double)r->cost : 1.0;
}
/* the same in a batch of mixed lengths (the slice dealing's costs) */
static double item_cost_mixed(int m, int v) {
  const pacg_batch_variant_t* r = variant_row(m, v);
  return r ? (double)r->cost_mixed : 1.0;
}

typedef struct {
  double cost;
  int idx;
} cost_idx_t;

static int by_cost_desc(const void* a, const void* b) {
  const cost_idx_t* x = (const cost_idx_t*)a;
  const cost_idx_t* y = (const cost_idx_t*)b;
  if (x->cost != y->cost) return x->cost > y->cost ? -1 : 1;
  return x->idx - y->idx;
}

/* Deal items (built longest first, one shape after another) to w workgroups:
 * longest processing time first, each to the least loaded workgroup; then
 * group them by workgroup, keeping the built order inside a group (so the
 * items of one shape stay together).  out[] receives the items, wg_first[]
 * the w + 1 group starts. */
static int deal_items(const pacg_batch_item_t* items, int n, int w, pacg_batch_item_t* out, uint32_t* wg_first) {
  cost_idx_t* ci = (cost_idx_t*)malloc((size_t)n * sizeof(cost_idx_t));
  int* owner = (int*)malloc((size_t)n * sizeof(int));
  double* load = (double*)calloc((size_t)w, sizeof(double));
  if (!ci || !owner || !load) {
    free(ci);
    free(owner);
    free(load);
    return pacg_fail(PACG_ERR_NOMEM, "out of host memory");
  }
  for (int i = 0; i < n; ++i) {
    ci[i].cost = item_cost(items[i].log2n, items[i].variant);
    ci[i].idx = i;
  }
  qsort(ci, (size_t)n, sizeof(cost_idx_t), by_cost_desc);
  for (int k = 0; k < n; ++k) {
    int best = 0;
    for (int j = 1; j < w; ++j)
      if (load[j] < load[best]) best = j;
    load[best] += ci[k].cost;
    owner[ci[k].idx] = best;
  }
  for (int j = 0; j <= w; ++j) wg_first[j] = 0;
  for (int i = 0; i < n; ++i) wg_first[owner[i] + 1]++;
  for (int j = 0; j < w; ++j) wg_first[j + 1] += wg_first[j];
  /* counting sort by owner, stable in the built order */
  uint32_t* at = (uint32_t*)malloc((size_t)w * sizeof(uint32_t));
  if (!at) {
    free(ci);
    free(owner);
    free(load);
    return pacg_fail(PACG_ERR_NOMEM, "out of host memory");
  }
  memcpy(at, wg_first, (size_t)w * sizeof(uint32_t));
  for (int i = 0; i < n; ++i) out[at[owner[i]]++] = items[i];
  free(at);
  free(ci);
  free(owner);
  free(load);
  return PACG_OK;
}

/* Time (0.1 us) a workgroup spends each time its list starts a new shape: the
 * run's first k is loaded before its loop, not behind the previous item's
 * phases (fitted with the mixed-batch item costs). */
#define RUN_COST 36.0

/* The slice dealing: items (built longest first, one shape after another) go
 * to w workgroups as contiguous slices of the built order, so that a workgroup
 * meets one shape, or two at a slice boundary.  Workgroup j takes the items
 * whose running mixed-batch cost (plus RUN_COST per slice) falls into its
 * share j / w .. (j + 1) / w of the total, judged at each item's middle; every
 * workgroup gets an item.  Balanced only as well as the costs are right, so
 * the autotuner keeps it only where it measures faster. */
static int deal_slices(const pacg_batch_item_t* items, int n, int w, pacg_batch_item_t* out, uint32_t* wg_first) {
  double total = RUN_COST * w;
  for (int i = 0; i < n; ++i) total += item_cost_mixed(items[i].log2n, items[i].variant);
  const double share = total / w;
  double acc = RUN_COST;
  int j = 0;
  for (int k = 0; k <= w; ++k) wg_first[k] = 0;
  for (int i = 0; i < n; ++i) {
    const double c = item_cost_mixed(items[i].log2n, items[i].variant);
    /* next workgroup once this one has its share, or when the items left are
     * only enough for one per workgroup left */
    if (j < w - 1 && i > 0 && (acc + 0.5 * c > (j + 1) * share || n - i <= w - 1 - j)) {
      ++j;
      acc += RUN_COST;
    }
    wg_first[j + 1]++;
    acc += c;
    out[i] = items[i];
  }
  for (int k = 0; k < w; ++k) wg_first[k + 1] += wg_first[k];
  return PACG_OK;
}

/* items for the lengths whose mask bit is set, dealt to the workgroups; uploads them */
static int build_items(pacg_batch* b, unsigned mask) {
  int n = 0;
  pacg_batch_item_t* items = (pacg_batch_item_t*)malloc((size_t)(b->count + 1) * sizeof(pacg_batch_item_t));
  pacg_batch_item_t* dealt = (pacg_batch_item_t*)malloc((size_t)(b->count + 1) * sizeof(pacg_batch_item_t));
  uint32_t* wg_first = (uint32_t*)malloc((size_t)(b->blocks + 1) * sizeof(uint32_t));
  if (!items || !dealt || !wg_first) {
    free(items);
    free(dealt);
    free(wg_first);
    return pacg_fail(PACG_ERR_NOMEM, "out of host memory");
  }
  int pos = 0;
  while (pos < b->count) {
    int m = b->log2n[b->order[pos]];
    int end = pos;
    while (end < b->count && b->log2n[b->order[end]] == m) ++end;
    if (mask & (1u << m)) {
      int g = per_item(m, b->variant[m]);
      for (int f = pos; f < end; f += g) {
        items[n].first = (uint32_t)f;
        items[n].count = (uint16_t)((end - f) < g ? (end - f) : g);
        items[n].log2n = (uint8_t)m;
        items[n].variant = (uint8_t)b->variant[m];
        ++n;
      }
    }
    pos = end;
  }
  int w = n < b->blocks ? n : b->blocks;
  int rc = PACG_OK;
  if (n > 0)
    rc = b->dealing == PACG_DEAL_SLICES ? deal_slices(items, n, w, dealt, wg_first)
                                        : deal_items(items, n, w, dealt, wg_first);
  if (!rc && n > 0 &&
      (pacg_memcpy_h2d(b->d_items, dealt, (size_t)n * sizeof(pacg_batch_item_t), b->stream) ||
       pacg_memcpy_h2d(b->d_wg_first, wg_first, (size_t)(w + 1) * sizeof(uint32_t), b->stream)))
    rc = pacg_fail(PACG_ERR_HIP, "item upload failed");
  if (!rc && pacg_stream_sync(b->stream)) rc = pacg_fail(PACG_ERR_HIP, "item upload failed");
  free(items);
  free(dealt);
  free(wg_first);
  b->n_items = rc ? 0 : n;
  b->n_wg = rc ? 0 : w;
  return rc;
}

/* order[] <- pair indices sorted by length, longest first, input order within a
 * length (a counting sort over the few possible lengths: no shared state, so
 * batches may be created from several threads at once) */
static void sort_longest_first(const int* log2n, int count, int* order) {
  int start[LMAX + 2] = {0};
  for (int i = 0; i < count; ++i) start[LMAX - log2n[i] + 1]++;
  for (int m = 1; m <= LMAX + 1; ++m) start[m] += start[m - 1];
  for (int i = 0; i < count; ++i) order[start[LMAX - log2n[i]]++] = i;
}

int pacg_batch_destroy(pacg_batch* b) {
  if (!b) return PACG_OK;
  if (b->stream) pacg_stream_sync(b->stream);
  if (b->graph) pacg_graph_destroy(b->graph);
  pacg_dev_free(b->d_conv_off);
  pacg_dev_free(b->d_items);
  pacg_dev_free(b->d_wg_first);
  pacg_dev_free(b->d_tw);
  if (b->own_stream && b->stream) pacg_stream_destroy(b->stream);
  free(b->log2n);
  free(b->offset);
  free(b->order);
  free(b);
  return PACG_OK;
}

int pacg_batch_create(pacg_batch** out, int count, const int* log2_len) {
  if (!out || count < 1 || !log2_len) return pacg_fail(PACG_ERR_ARG, "pacg_batch_create: bad argument");
  *out = NULL;
  pacg_batch* b = (pacg_batch*)calloc(1, sizeof *b);
  if (!b) return pacg_fail(PACG_ERR_NOMEM, "out of host memory");
  b->count = count;
  b->log2n = (int*)malloc((size_t)count * sizeof(int));
  b->offset = (long long*)malloc((size_t)count * sizeof(long long));
  b->order = (int*)malloc((size_t)count * sizeof(int));
  if (!b->log2n || !b->offset || !b->order) {
    pacg_batch_destroy(b);
    return pacg_fail(PACG_ERR_NOMEM, "out of host memory");
  }
  long long off = 0;
  for (int i = 0; i < count; ++i) {
    int m = log2_len[i];
    if (m < LMIN || m > LMAX) {
      pacg_batch_destroy(b);
      return pacg_fail(PACG_ERR_SIZE, "pair %d: length 2^%d outside 2^%d..2^%d", i, m, LMIN, LMAX);
    }
    b->log2n[i] = m;
    b->offset[i] = off;
    off += 1LL << m;
  }
  b->total = off;
  if (off >= (1LL << 30)) { /* the kernel addresses each array with 32-bit byte offsets */
    pacg_batch_destroy(b);
    return pacg_fail(PACG_ERR_SIZE, "batch of %lld points reaches 2^30", off);
  }
  sort_longest_first(b->log2n, count, b->order);
  for (int m = 0; m <= LMAX; ++m) b->variant[m] = 0;
  int rc = PACG_OK, cus = 0;
  if (pacg_stream_create(&b->stream)) {
    rc = pacg_fail(PACG_ERR_HIP, "stream creation failed");
    goto fail;
  }
  b->own_stream = 1;
  if (pacg_cu_count(&cus) || cus < 1) cus = 256;
  b->blocks = cus;
  b->cap_items = count + 1;
  if (pacg_dev_malloc((void**)&b->d_conv_off, (size_t)count * sizeof(uint32_t)) ||
      pacg_dev_malloc((void**)&b->d_items, (size_t)b->cap_items * sizeof(pacg_batch_item_t)) ||
      pacg_dev_malloc((void**)&b->d_wg_first, (size_t)(b->blocks + 1) * sizeof(uint32_t)) ||
      pacg_dev_malloc(&b->d_tw, (size_t)ITEM_POINTS * 2 * sizeof(float))) {
    rc = pacg_fail(PACG_ERR_NOMEM, "device allocation failed");
    goto fail;
  }
  {
    uint32_t* co = (uint32_t*)malloc((size_t)count * sizeof(uint32_t));
    if (!co) {
      rc = pacg_fail(PACG_ERR_NOMEM, "out of host memory");
      goto fail;
    }
    for (int p = 0; p < count; ++p) co[p] = (uint32_t)b->offset[b->order[p]];
    int e = pacg_memcpy_h2d(b->d_conv_off, co, (size_t)count * sizeof(uint32_t), b->stream);
    e |= pacg_stream_sync(b->stream);
    free(co);
    if (e) {
      rc = pacg_fail(PACG_ERR_HIP, "offset upload failed");
      goto fail;
    }
  }
  /* w_{2^14} in FP64 on the device, rounded to FP32 (twiddle_factors.c:17-131 values) */
  if (pacg_gen_twiddles(b->d_tw, PACG_F32, ITEM_POINTS, ITEM_POINTS, 1, b->stream)) {
    rc = pacg_fail(PACG_ERR_HIP, "twiddle generation failed");
    goto fail;
  }
  if ((rc = build_items(b, ~0u))) goto fail;
  *out = b;
  pacg_clear_error();
  return PACG_OK;
fail:
  pacg_batch_destroy(b);
  return rc;
}

long long pacg_batch_total(const pacg_batch* b) { return b ? b->total : 0; }
long long pacg_batch_offset(const pacg_batch* b, int i) { return (b && i >= 0 && i < b->count) ? b->offset[i] : -1; }
void* pacg_batch_stream(const pacg_batch* b) { return b ? b->stream : NULL; }
int pacg_batch_items(const pacg_batch* b) { return b ? b->n_items : 0; }

int pacg_batch_get_variant(const pacg_batch* b, int log2n) {
  if (!b || log2n < LMIN || log2n > LMAX) return -1;
  return b->variant[log2n];
}

int pacg_batch_set_dealing(pacg_batch* b, int dealing) {
  if (!b || (dealing != PACG_DEAL_LPT && dealing != PACG_DEAL_SLICES))
    return pacg_fail(PACG_ERR_ARG, "pacg_batch_set_dealing: bad argument");
  b->dealing = dealing;
  int rc = build_items(b, ~0u);
  if (b->graph) {
    pacg_graph_destroy(b->graph);
    b->graph = NULL;
  }
  return rc;
}

int pacg_batch_get_dealing(const pacg_batch* b) {
  if (!b) return -pacg_fail(PACG_ERR_ARG, "pacg_batch_get_dealing: NULL batch");
  return b->dealing;
}

int pacg_batch_set_variant(pacg_batch* b, int log2n, int variant) {
  if (!b || log2n < LMIN || log2n > LMAX || !variant_row(log2n, variant))
    return pacg_fail(PACG_ERR_ARG, "no batched variant %d for length 2^%d", variant, log2n);
  b->variant[log2n] = variant;
  if (b->graph) {
    pacg_graph_destroy(b->graph);
    b->graph = NULL;
  }
  return build_items(b, ~0u);
}

static int launch(pacg_batch* b, const float* xr, const float* xi, const float* kr, const float* ki, float* yr,
                  float* yi) {
  pacg_batch_args_t a;
  memset(&a, 0, sizeof a);
  a.xr = xr;
  a.xi = xi;
  a.kr = kr;
  a.ki = ki;
  a.yr = yr;
  a.yi = yi;
  a.conv_off = b->d_conv_off;
  a.items = b->d_items;
  a.wg_first = b->d_wg_first;
  a.n_items = (uint32_t)b->n_items;
  a.tw = b->d_tw;
  a.total = (uint32_t)b->total;
  int rc = pacg_batch_launch(&a, b->n_wg, b->stream);
  if (rc) return pacg_fail(PACG_ERR_HIP, "batched launch failed: %s", pacg_hip_error_string(rc));
  return PACG_OK;
}

int pacg_batch_convolve(pacg_batch* b, const float* xr, const float* xi, const float* kr, const float* ki, float* yr,
                        float* yi) {
  if (!b || !xr || !xi || !kr || !ki || !yr || !yi) return pacg_fail(PACG_ERR_ARG, "pacg_batch_convolve: NULL");
  return launch(b, xr, xi, kr, ki, yr, yi);
}

/* whole-batch launch time in ms: one warm launch, then the mean of `reps` */
static int time_batch(pacg_batch* b, const float* xr, const float* xi, const float* kr, const float* ki, float* yr,
                      float* yi, int reps, void* e0, void* e1, float* ms) {
  int rc = launch(b, xr, xi, kr, ki, yr, yi);
  if (rc) return rc;
  pacg_event_record(e0, b->stream);
  for (int r = 0; r < reps && !rc; ++r) rc = launch(b, xr, xi, kr, ki, yr, yi);
  pacg_event_record(e1, b->stream);
  if (pacg_stream_sync(b->stream)) return pacg_fail(PACG_ERR_HIP, "autotune timing failed");
  if (rc) return rc;
  *ms = 0.0f;
  pacg_event_elapsed_ms(ms, e0, e1);
  *ms /= (float)reps;
  return PACG_OK;
}

/* Coordinate descent over the lengths present, longest first: each variant of
 * a length is timed on the whole batch (the other lengths at their current
 * variants), because what a variant costs depends on how its items fill the
 * workgroups beside the others.  Two sweeps, each variant keeping its faster
 * measurement; a length keeps its current variant unless another is faster by
 * more than 1 %. */
/* Coordinate descent over the variants of the lengths present, with the
 * batch's current dealing: two sweeps, each variant timed twice on the whole
 * batch (its faster timing kept), a variant replacing the current one only
 * when more than 1 % faster. */
static int sweep_variants(pacg_batch* b, const float* xr, const float* xi, const float* kr, const float* ki, float* yr,
                          float* yi, int reps, void* e0, void* e1) {
  int rc = PACG_OK;
  for (int sweep = 0; sweep < 2 && !rc; ++sweep) {
    for (int m = LMAX; m >= LMIN && !rc; --m) {
      int present = 0;
      for (int i = 0; i < b->count && !present; ++i) present = b->log2n[i] == m;
      const int nv = pacg_batch_num_variants(m);
      if (!present || nv < 2) continue;
      enum { MAXV = 8 };
      float vms[MAXV];
      const int cur = b->variant[m];
      for (int v = 0; v < nv && v < MAXV && !rc; ++v) {
        b->variant[m] = v;
        if ((rc = build_items(b, ~0u))) break;
        float t1 = 0.0f, t2 = 0.0f;
        if ((rc = time_batch(b, xr, xi, kr, ki, yr, yi, reps, e0, e1, &t1))) break;
        if ((rc = time_batch(b, xr, xi, kr, ki, yr, yi, reps, e0, e1, &t2))) break;
        vms[v] = t1 < t2 ? t1 : t2;
      }
      if (rc) break;
      int best = cur;
      for (int v = 0; v < nv && v < MAXV; ++v)
        if (vms[v] < vms[best]) best = v;
      if (vms[cur] <= 1.01f * vms[best]) best = cur;
      b->variant[m] = best;
    }
  }
  return rc;
}

/* The variants for each dealing in turn (sweeps from the batch's starting
 * variants), each result timed four times (its fastest kept); the slices
 * replace the default dealing, longest processing time first, only when more
 * than 1 % faster.  The dealing and the variants interact: a variant with
 * smaller items balances the default dealing, the slices need fewer shapes
 * per workgroup, so they are searched together. */
int pacg_batch_autotune(pacg_batch* b, const float* xr, const float* xi, const float* kr, const float* ki, float* yr,
                        float* yi, int reps) {
  if (!b || !xr || !xi || !kr || !ki || !yr || !yi || reps < 1) return -pacg_fail(PACG_ERR_ARG, "bad argument");
  void *e0 = NULL, *e1 = NULL;
  if (pacg_event_create(&e0) || pacg_event_create(&e1)) {
    if (e0) pacg_event_destroy(e0);
    return -pacg_fail(PACG_ERR_HIP, "event creation failed");
  }
  int start[LMAX + 1], found[2][LMAX + 1];
  for (int m = 0; m <= LMAX; ++m) start[m] = b->variant[m];
  const int had = b->dealing;
  float dms[2] = {1e30f, 1e30f};
  int rc = PACG_OK;
  for (int d = 0; d < 2 && !rc; ++d) {
    b->dealing = d;
    for (int m = 0; m <= LMAX; ++m) b->variant[m] = start[m];
    if ((rc = sweep_variants(b, xr, xi, kr, ki, yr, yi, reps, e0, e1))) break;
    for (int m = 0; m <= LMAX; ++m) found[d][m] = b->variant[m];
    if ((rc = build_items(b, ~0u))) break;
    for (int t = 0; t < 4 && !rc; ++t) {
      float ms = 0.0f;
      if ((rc = time_batch(b, xr, xi, kr, ki, yr, yi, reps, e0, e1, &ms))) break;
      if (ms < dms[d]) dms[d] = ms;
    }
  }
  pacg_event_destroy(e0);
  pacg_event_destroy(e1);
  if (rc) { /* leave the batch as it was */
    b->dealing = had;
    for (int m = 0; m <= LMAX; ++m) b->variant[m] = start[m];
  } else {
    b->dealing = dms[PACG_DEAL_SLICES] < 0.99f * dms[PACG_DEAL_LPT] ? PACG_DEAL_SLICES : PACG_DEAL_LPT;
    for (int m = 0; m <= LMAX; ++m) b->variant[m] = found[b->dealing][m];
  }
  int rc2 = build_items(b, ~0u);
  if (b->graph) {
    pacg_graph_destroy(b->graph);
    b->graph = NULL;
  }
  if (rc) return -rc;
  if (rc2) return -rc2;
  int changed = 0;
  for (int m = LMIN; m <= LMAX; ++m) changed += b->variant[m] != start[m];
  return changed;
}

int pacg_batch_capture(pacg_batch* b, const float* xr, const float* xi, const float* kr, const float* ki, float* yr,
                       float* yi) {
  if (!b) return pacg_fail(PACG_ERR_ARG, "NULL batch");
  if (b->graph) {
    pacg_graph_destroy(b->graph);
    b->graph = NULL;
  }
  if (pacg_graph_begin(b->stream)) return pacg_fail(PACG_ERR_HIP, "begin capture failed");
  int rc = launch(b, xr, xi, kr, ki, yr, yi);
  void* exec = NULL;
  int rc2 = pacg_graph_end(b->stream, &exec);
  if (rc) return rc;
  if (rc2) return pacg_fail(PACG_ERR_HIP, "end capture failed: %s", pacg_hip_error_string(rc2));
  b->graph = exec;
  return PACG_OK;
}

int pacg_batch_replay(pacg_batch* b) {
  if (!b || !b->graph) return pacg_fail(PACG_ERR_STATE, "no captured batch graph");
  if (pacg_graph_launch(b->graph, b->stream)) return pacg_fail(PACG_ERR_HIP, "graph launch failed");
  return PACG_OK;
}

int pacg_batch_sync(pacg_batch* b) {
  if (!b) return pacg_fail(PACG_ERR_ARG, "NULL batch");
  if (pacg_stream_sync(b->stream)) return pacg_fail(PACG_ERR_HIP, "stream synchronize failed");
  return PACG_OK;
}
