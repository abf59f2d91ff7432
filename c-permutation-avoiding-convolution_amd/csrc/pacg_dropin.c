/*
 * Drop-in GPU entry points with the reference's fft.h names (fft.h:24-221).
 *
 * Each call validates its arguments, takes (or builds) the cached FP64 plan
 * of its shape and radix, copies the host arrays to the device, runs the
 * pass kernels and copies every array the reference modifies back:
 *   apply_*butterflies_*      re/im                  (butterflies*.c)
 *   fft / ifft / fft2 / u*    re/im                  (fft*.c)
 *   permutation_avoiding_*    re/im and h <- A^T h   (convolution.c:45-53)
 *   fft_based_convolution*    re/im and h <- F h     (convolution.c:18)
 * The caller's twiddle tables are not read (the device builds its own FP64
 * tables); the caller's rho is uploaded and applied as a gather.
 * On invalid input nothing is modified and pacg_last_error() says why.
 * Calls are serialised by one mutex; each is synchronous.  Host arrays move
 * through pinned memory: a staging buffer per cached shape up to 1 MiB per
 * array, a shared chunked ring (host copies on a small thread pool) above.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "pacg.h"
#include "pacg_fft.h"
#include "pacg_host.h"
#include "pacg_internal.h"

#define NSLOT 8
/* Arrays up to this size travel through a pinned staging buffer of the slot
 * (host memcpy + DMA, both directions): the runtime's pageable path costs
 * about 180 us for a 512 KiB host-to-device copy on MI355X, and its calls now
 * and then take several times longer; staged, about 30 us.  Larger arrays take
 * the pageable path, which streams at 30 GB/s (host to device). */
#define STAGE_MAX_BYTES (1u << 20)

typedef struct {
  int nd, dims[3], r;
  pacg_plan* plan;
  void *xr, *xi, *hr, *hi, *sr, *si;
  int* rho;
  double* stage; /* pinned, 4 arrays (x re/im, h re/im), or NULL */
  unsigned long long used;
} slot_t;

static slot_t g_slot[NSLOT];
static unsigned long long g_tick;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

static void slot_free(slot_t* s) {
  pacg_plan_destroy(s->plan);
  pacg_dev_free(s->xr);
  pacg_dev_free(s->xi);
  pacg_dev_free(s->hr);
  pacg_dev_free(s->hi);
  pacg_dev_free(s->sr);
  pacg_dev_free(s->si);
  pacg_dev_free(s->rho);
  pacg_host_free(s->stage);
  memset(s, 0, sizeof *s);
}

/* radix r valid for the reference: n = r^t exactly per axis */
static int check_shape(int nd, const int* dims, int r) {
  if (r != 2 && r != 4 && r != 8) return pacg_fail(PACG_ERR_ARG, "radix %d not in {2,4,8}", r);
  long long n = 1;
  for (int a = 0; a < nd; ++a) {
    if (dims[a] < 1) return pacg_fail(PACG_ERR_SIZE, "axis %d length %d", a, dims[a]);
    long long m = 1;
    while (m < dims[a]) m *= r;
    if (m != dims[a]) return pacg_fail(PACG_ERR_SIZE, "axis %d length %d is not a power of %d", a, dims[a], r);
    n *= dims[a];
  }
  if (n >= (1LL << 31)) return pacg_fail(PACG_ERR_SIZE, "total size %lld >= 2^31", n);
  return PACG_OK;
}

static slot_t* get_slot(int nd, const int* dims, int r) {
  for (int i = 0; i < NSLOT; ++i) {
    slot_t* s = &g_slot[i];
    if (s->plan && s->nd == nd && s->r == r && s->dims[0] == dims[0] && (nd < 2 || s->dims[1] == dims[1]) &&
        (nd < 3 || s->dims[2] == dims[2])) {
      s->used = ++g_tick;
      return s;
    }
  }
  slot_t* victim = &g_slot[0];
  for (int i = 0; i < NSLOT; ++i) {
    if (!g_slot[i].plan) {
      victim = &g_slot[i];
      break;
    }
    if (g_slot[i].used < victim->used) victim = &g_slot[i];
  }
  if (victim->plan) slot_free(victim);
  pacg_desc d;
  if (pacg_desc_init(&d, nd, dims, r, PACG_PREC_F64)) return NULL;
  if (pacg_plan_create(&victim->plan, &d)) return NULL;
  size_t bytes = (size_t)pacg_plan_size(victim->plan) * sizeof(double);
  if (pacg_dev_malloc(&victim->xr, bytes) || pacg_dev_malloc(&victim->xi, bytes) ||
      pacg_dev_malloc(&victim->hr, bytes) || pacg_dev_malloc(&victim->hi, bytes) ||
      pacg_dev_malloc(&victim->sr, bytes) || pacg_dev_malloc(&victim->si, bytes) ||
      pacg_dev_malloc((void**)&victim->rho, (size_t)pacg_plan_size(victim->plan) * sizeof(int))) {
    slot_free(victim);
    pacg_fail(PACG_ERR_NOMEM, "device allocation for a %lld-element plan failed", (long long)(bytes / 8));
    return NULL;
  }
  /* the staging buffer is an optimisation: without it the copies take the pageable path */
  if (bytes <= STAGE_MAX_BYTES && pacg_host_malloc((void**)&victim->stage, 4 * bytes)) victim->stage = NULL;
  victim->nd = nd;
  victim->r = r;
  for (int a = 0; a < 3; ++a) victim->dims[a] = a < nd ? dims[a] : 0;
  victim->used = ++g_tick;
  return victim;
}

/* ---- large arrays: pinned ring, chunked and pipelined ---------------------
 * The runtime's pageable path streams at about 30 GB/s host to device and
 * 17 GB/s device to host on MI355X; DMA from pinned memory runs at 56 GB/s.
 * Arrays above STAGE_MAX_BYTES therefore go through RING_SLOTS pinned chunks:
 * the host copy of one chunk (split over a small thread pool, one host thread
 * alone copies too slowly) overlaps the DMA of the previous ones.  The ring is
 * shared by all slots; every drop-in call holds g_mu while it uses it. */
#define RING_CHUNK ((size_t)8 << 20)
#define RING_SLOTS 3
#define POOL_MAX 8
#define PART_MIN ((size_t)1 << 20)

static char* g_ring[RING_SLOTS];
static void* g_ring_ev[RING_SLOTS];
static int g_ring_state; /* 0 not tried, 1 ready, -1 unavailable (pageable path) */

static pthread_t g_pool[POOL_MAX - 1];
static int g_pool_n = -1; /* worker threads, -1 before the first use */
static pthread_mutex_t g_pool_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t g_pool_go = PTHREAD_COND_INITIALIZER;
static pthread_cond_t g_pool_done = PTHREAD_COND_INITIALIZER;
static struct {
  char* dst;
  const char* src;
  size_t bytes, part;
  int parts, next, done;
  unsigned gen;
} g_job;

/* take parts of the current job until none is left (g_pool_mu held on entry and exit) */
static void pool_work(void) {
  while (g_job.next < g_job.parts) {
    const int i = g_job.next++;
    const size_t off = (size_t)i * g_job.part;
    const size_t len = off + g_job.part <= g_job.bytes ? g_job.part : g_job.bytes - off;
    pthread_mutex_unlock(&g_pool_mu);
    memcpy(g_job.dst + off, g_job.src + off, len);
    pthread_mutex_lock(&g_pool_mu);
    if (++g_job.done == g_job.parts) pthread_cond_broadcast(&g_pool_done);
  }
}

static void* pool_main(void* arg) {
  (void)arg;
  unsigned seen = 0;
  pthread_mutex_lock(&g_pool_mu);
  for (;;) {
    while (g_job.gen == seen) pthread_cond_wait(&g_pool_go, &g_pool_mu);
    seen = g_job.gen;
    pool_work();
  }
  return NULL;
}

static void pool_start(void) {
  long cpus = sysconf(_SC_NPROCESSORS_ONLN);
  int want = cpus > 1 ? (int)(cpus < POOL_MAX ? cpus : POOL_MAX) - 1 : 0;
  g_pool_n = 0;
  for (int i = 0; i < want; ++i) {
    if (pthread_create(&g_pool[i], NULL, pool_main, NULL)) break;
    pthread_detach(g_pool[i]);
    g_pool_n++;
  }
}

/* memcpy split over the pool; the calling thread takes parts too */
static void par_memcpy(void* dst, const void* src, size_t bytes) {
  if (g_pool_n < 0) pool_start();
  if (g_pool_n == 0 || bytes < 2 * PART_MIN) {
    memcpy(dst, src, bytes);
    return;
  }
  int parts = g_pool_n + 1;
  size_t part = (bytes + (size_t)parts - 1) / (size_t)parts;
  if (part < PART_MIN) part = PART_MIN;
  pthread_mutex_lock(&g_pool_mu);
  g_job.dst = (char*)dst;
  g_job.src = (const char*)src;
  g_job.bytes = bytes;
  g_job.part = part;
  g_job.parts = (int)((bytes + part - 1) / part);
  g_job.next = 0;
  g_job.done = 0;
  g_job.gen++;
  pthread_cond_broadcast(&g_pool_go);
  pool_work();
  while (g_job.done < g_job.parts) pthread_cond_wait(&g_pool_done, &g_pool_mu);
  pthread_mutex_unlock(&g_pool_mu);
}

static int ring_ready(void) {
  if (g_ring_state == 0) {
    g_ring_state = 1;
    for (int i = 0; i < RING_SLOTS; ++i)
      if (pacg_host_malloc((void**)&g_ring[i], RING_CHUNK) || pacg_event_create(&g_ring_ev[i])) g_ring_state = -1;
  }
  return g_ring_state == 1;
}

static void ring_release(void) {
  for (int i = 0; i < RING_SLOTS; ++i) {
    if (g_ring_ev[i]) pacg_event_sync(g_ring_ev[i]);
    pacg_host_free(g_ring[i]);
    if (g_ring_ev[i]) pacg_event_destroy(g_ring_ev[i]);
    g_ring[i] = NULL;
    g_ring_ev[i] = NULL;
  }
  g_ring_state = 0;
}

/* chunk c of every transfer uses ring slot c % RING_SLOTS; g_ring_used[i]
 * says whether slot i has a DMA recorded that a new user must wait for */
static int g_ring_used[RING_SLOTS];

static int ring_h2d(void* d, const double* h, size_t bytes, void* stream) {
  int e = 0;
  size_t nch = (bytes + RING_CHUNK - 1) / RING_CHUNK;
  for (size_t c = 0; c < nch && !e; ++c) {
    const int k = (int)(c % RING_SLOTS);
    const size_t off = c * RING_CHUNK, len = off + RING_CHUNK <= bytes ? RING_CHUNK : bytes - off;
    if (g_ring_used[k]) e |= pacg_event_sync(g_ring_ev[k]);  /* its previous DMA has read it */
    if (e) break;
    par_memcpy(g_ring[k], (const char*)h + off, len);
    e |= pacg_memcpy_h2d((char*)d + off, g_ring[k], len, stream);
    e |= pacg_event_record(g_ring_ev[k], stream);
    g_ring_used[k] = 1;
  }
  return e;
}

static int ring_d2h(double* h, const void* d, size_t bytes, void* stream) {
  int e = 0;
  size_t nch = (bytes + RING_CHUNK - 1) / RING_CHUNK;
  /* keep RING_SLOTS - 1 DMAs ahead of the host copies */
  for (size_t c = 0; c < nch + RING_SLOTS - 1 && !e; ++c) {
    if (c < nch) {
      const int k = (int)(c % RING_SLOTS);
      const size_t off = c * RING_CHUNK, len = off + RING_CHUNK <= bytes ? RING_CHUNK : bytes - off;
      /* slot k's previous chunk was copied out one iteration ago (host side), and any
       * host-to-device DMA still reading it is ordered before this copy on the stream */
      e |= pacg_memcpy_d2h(g_ring[k], (const char*)d + off, len, stream);
      e |= pacg_event_record(g_ring_ev[k], stream);
      g_ring_used[k] = 1;
    }
    if (c + 1 >= RING_SLOTS && c + 1 - RING_SLOTS < nch) {
      const size_t j = c + 1 - RING_SLOTS;
      const int k = (int)(j % RING_SLOTS);
      const size_t off = j * RING_CHUNK, len = off + RING_CHUNK <= bytes ? RING_CHUNK : bytes - off;
      e |= pacg_event_sync(g_ring_ev[k]);
      if (!e) par_memcpy((char*)h + off, g_ring[k], len);
    }
  }
  return e;
}

/* host array h -> device array d; k (0..3) picks the array's region of the staging buffer */
static int up(slot_t* s, void* d, const double* h, int k) {
  const size_t n = (size_t)pacg_plan_size(s->plan);
  if (s->stage) {
    double* st = s->stage + (size_t)k * n;
    memcpy(st, h, n * sizeof(double));
    return pacg_memcpy_h2d(d, st, n * sizeof(double), pacg_plan_stream(s->plan));
  }
  if (ring_ready()) return ring_h2d(d, h, n * sizeof(double), pacg_plan_stream(s->plan));
  return pacg_memcpy_h2d(d, h, n * sizeof(double), pacg_plan_stream(s->plan));
}
/* device array d -> host array h; staged slots land in region k of the staging
 * buffer and are copied out by finish_down() after the stream is synchronised */
static int down(slot_t* s, double* h, const void* d, int k) {
  const size_t n = (size_t)pacg_plan_size(s->plan);
  if (s->stage) return pacg_memcpy_d2h(s->stage + (size_t)k * n, d, n * sizeof(double), pacg_plan_stream(s->plan));
  if (ring_ready()) return ring_d2h(h, d, n * sizeof(double), pacg_plan_stream(s->plan));
  return pacg_memcpy_d2h(h, d, n * sizeof(double), pacg_plan_stream(s->plan));
}
static void finish_down(slot_t* s, double* h, int k) {
  const size_t n = (size_t)pacg_plan_size(s->plan);
  if (s->stage) memcpy(h, s->stage + (size_t)k * n, n * sizeof(double));
}

enum { OPK_TRANS, OPK_CONJ, OPK_FWD, OPK_FFT, OPK_FFT2, OPK_UFFT, OPK_IFFT, OPK_UIFFT, OPK_PA, OPK_FB };

/* One drop-in call: x (and h for the convolutions) are host arrays. */
static void dropin(int op, int nd, const int* dims, int r, double* re, double* im, const int* rho, double* hre,
                   double* him) {
  if (!re || !im || ((op == OPK_PA || op == OPK_FB) && (!hre || !him)) ||
      ((op == OPK_FFT || op == OPK_FFT2 || op == OPK_IFFT || op == OPK_FB) && !rho)) {
    pacg_fail(PACG_ERR_ARG, "NULL array argument");
    return;
  }
  if (check_shape(nd, dims, r)) return;
  if (rho) {
    long long n = 1;
    for (int a = 0; a < nd; ++a) n *= dims[a];
    for (long long j = 0; j < n; ++j)
      if (rho[j] < 0 || rho[j] >= n) {
        pacg_fail(PACG_ERR_ARG, "rho[%lld] = %d outside [0, %lld)", j, rho[j], n);
        return;
      }
  }
  pthread_mutex_lock(&g_mu);
  slot_t* s = get_slot(nd, dims, r);
  if (!s) {
    pthread_mutex_unlock(&g_mu);
    return;
  }
  pacg_plan* pl = s->plan;
  void* st = pacg_plan_stream(pl);
  long long n = pacg_plan_size(pl);
  int rc = 0, ok = 0;
  rc |= up(s, s->xr, re, 0);
  rc |= up(s, s->xi, im, 1);
  if (op == OPK_PA || op == OPK_FB) {
    rc |= up(s, s->hr, hre, 2);
    rc |= up(s, s->hi, him, 3);
  }
  if (rho) rc |= pacg_memcpy_h2d(s->rho, rho, (size_t)n * sizeof(int), st);
  if (rc) {
    pacg_fail(PACG_ERR_HIP, "host to device copy failed");
    goto done;
  }
  switch (op) {
    case OPK_TRANS:
      rc = pacg_execute(pl, PACG_OP_TRANS, s->xr, s->xi, s->xr, s->xi);
      break;
    case OPK_CONJ:
      rc = pacg_execute(pl, PACG_OP_CONJ, s->xr, s->xi, s->xr, s->xi);
      break;
    case OPK_FWD:
    case OPK_UFFT:
      rc = pacg_execute(pl, PACG_OP_FWD, s->xr, s->xi, s->xr, s->xi);
      break;
    case OPK_UIFFT:
      rc = pacg_execute(pl, PACG_OP_UIFFT, s->xr, s->xi, s->xr, s->xi);
      break;
    case OPK_FFT: /* fft.c:12-25: P then A */
      rc = pacg_gather_rho(s->xr, s->xi, s->sr, s->si, s->rho, (uint64_t)n, PACG_F64, st);
      if (!rc) rc = pacg_execute(pl, PACG_OP_FWD, s->sr, s->si, s->xr, s->xi);
      break;
    case OPK_IFFT: /* fft.c:55-71: P, conj A, /n */
      rc = pacg_gather_rho(s->xr, s->xi, s->sr, s->si, s->rho, (uint64_t)n, PACG_F64, st);
      if (!rc) rc = pacg_execute(pl, PACG_OP_UIFFT, s->sr, s->si, s->xr, s->xi);
      break;
    case OPK_FFT2: /* fft.c:27-40: A^T then P */
      rc = pacg_execute(pl, PACG_OP_TRANS, s->xr, s->xi, s->sr, s->si);
      if (!rc) rc = pacg_gather_rho(s->sr, s->si, s->xr, s->xi, s->rho, (uint64_t)n, PACG_F64, st);
      break;
    case OPK_PA: /* convolution.c:31-72 */
      rc = pacg_convolve_pair(pl, s->xr, s->xi, s->hr, s->hi);
      break;
    case OPK_FB: /* convolution.c:12-29 with the caller's rho */
      rc = pacg_gather_rho(s->xr, s->xi, s->sr, s->si, s->rho, (uint64_t)n, PACG_F64, st);
      if (!rc) rc = pacg_execute(pl, PACG_OP_FWD, s->sr, s->si, s->xr, s->xi);
      if (!rc) rc = pacg_gather_rho(s->hr, s->hi, s->sr, s->si, s->rho, (uint64_t)n, PACG_F64, st);
      if (!rc) rc = pacg_execute(pl, PACG_OP_FWD, s->sr, s->si, s->hr, s->hi);
      if (!rc) rc = pacg_pointwise(s->xr, s->xi, s->hr, s->hi, PACG_F64, PACG_F64, (uint64_t)n, 1.0, st);
      if (!rc) rc = pacg_gather_rho(s->xr, s->xi, s->sr, s->si, s->rho, (uint64_t)n, PACG_F64, st);
      if (!rc) rc = pacg_execute(pl, PACG_OP_UIFFT, s->sr, s->si, s->xr, s->xi);
      break;
    default:
      rc = PACG_ERR_ARG;
  }
  if (rc) {
    if (!pacg_last_error()[0]) pacg_fail(PACG_ERR_HIP, "device execution failed (%d)", rc);
    goto done;
  }
  {
    /* Arrays up to 1 MiB come back through the slot's staging buffer and are
     * copied out only once every copy has succeeded, so a failure leaves the
     * caller's arrays untouched.  Larger arrays come back through the pinned
     * ring straight into the caller's arrays: a HIP failure during those copies
     * can leave them partly written. */
    int e = 0;
    e |= pacg_stream_sync(st);
    if (e) {
      pacg_fail(PACG_ERR_HIP, "device execution failed: %s", pacg_hip_error_string(e));
      goto done;
    }
    e |= down(s, re, s->xr, 0);
    e |= down(s, im, s->xi, 1);
    if (op == OPK_PA || op == OPK_FB) {
      e |= down(s, hre, s->hr, 2);
      e |= down(s, him, s->hi, 3);
    }
    e |= pacg_stream_sync(st);
    if (e) {
      pacg_fail(PACG_ERR_HIP, "device to host copy failed");
    } else {
      finish_down(s, re, 0);
      finish_down(s, im, 1);
      if (op == OPK_PA || op == OPK_FB) {
        finish_down(s, hre, 2);
        finish_down(s, him, 3);
      }
      pacg_clear_error();
      ok = 1;
    }
  }
done:
  /* on a failure, wait for the slot's stream before the next call may reuse the
   * staging buffers and the ring: no copy may still touch them or the caller's arrays */
  if (!ok) (void)pacg_stream_sync(st);
  pthread_mutex_unlock(&g_mu);
}

/* Release every cached drop-in plan and its device buffers. */
void pacg_dropin_release(void) {
  pthread_mutex_lock(&g_mu);
  for (int i = 0; i < NSLOT; ++i)
    if (g_slot[i].plan) slot_free(&g_slot[i]);
  ring_release();
  for (int i = 0; i < RING_SLOTS; ++i) g_ring_used[i] = 0;
  pthread_mutex_unlock(&g_mu);
}

#define D1(n) int dims_[1] = {n}
#define D2(a, b) int dims_[2] = {a, b}
#define D3(a, b, c) int dims_[3] = {a, b, c}

/* ---- butterflies (fft.h:24-99) ------------------------------------------- */
#define BFLY1(R)                                                                                                \
  void apply_butterflies_r##R(double* restrict re, double* restrict im, int n, const double** tr,              \
                              const double** ti) {                                                               \
    (void)tr, (void)ti;                                                                                         \
    D1(n);                                                                                                      \
    dropin(OPK_FWD, 1, dims_, R, re, im, NULL, NULL, NULL);                                                      \
  }                                                                                                             \
  void apply_conjugate_butterflies_r##R(double* restrict re, double* restrict im, int n, const double** tr,    \
                                        const double** ti) {                                                     \
    (void)tr, (void)ti;                                                                                         \
    D1(n);                                                                                                      \
    dropin(OPK_CONJ, 1, dims_, R, re, im, NULL, NULL, NULL);                                                     \
  }                                                                                                             \
  void apply_transposed_butterflies_r##R(double* restrict re, double* restrict im, int n, const double** tr,   \
                                         const double** ti) {                                                    \
    (void)tr, (void)ti;                                                                                         \
    D1(n);                                                                                                      \
    dropin(OPK_TRANS, 1, dims_, R, re, im, NULL, NULL, NULL);                                                    \
  }
BFLY1(2)
BFLY1(4)
BFLY1(8)

#define BFLY2(R, NAME, OP)                                                                                      \
  void NAME(double* restrict re, double* restrict im, int n1, int n2, const double** tr1, const double** ti1, \
            const double** tr2, const double** ti2) {                                                          \
    (void)tr1, (void)ti1, (void)tr2, (void)ti2;                                                                 \
    D2(n1, n2);                                                                                                 \
    dropin(OP, 2, dims_, R, re, im, NULL, NULL, NULL);                                                          \
  }
BFLY2(2, apply_butterflies_r2_2d, OPK_FWD)
BFLY2(2, apply_conjugate_butterflies_r2_2d, OPK_CONJ)
BFLY2(2, apply_transposed_butterflies_r2_2d, OPK_TRANS)
BFLY2(4, apply_butterflies_r4_2d, OPK_FWD)
BFLY2(4, apply_conjugate_butterflies_r4_2d, OPK_CONJ)
BFLY2(4, apply_transposed_butterflies_r4_2d, OPK_TRANS)

#define BFLY3(R, NAME, OP)                                                                                    \
  void NAME(double* restrict re, double* restrict im, int n1, int n2, int n3, const double** tr1,           \
            const double** ti1, const double** tr2, const double** ti2, const double** tr3,                  \
            const double** ti3) {                                                                            \
    (void)tr1, (void)ti1, (void)tr2, (void)ti2, (void)tr3, (void)ti3;                                         \
    D3(n1, n2, n3);                                                                                           \
    dropin(OP, 3, dims_, R, re, im, NULL, NULL, NULL);                                                        \
  }
BFLY3(2, apply_butterflies_r2_3d, OPK_FWD)
BFLY3(2, apply_conjugate_butterflies_r2_3d, OPK_CONJ)
BFLY3(2, apply_transposed_butterflies_r2_3d, OPK_TRANS)
BFLY3(4, apply_butterflies_r4_3d, OPK_FWD)
BFLY3(4, apply_conjugate_butterflies_r4_3d, OPK_CONJ)
BFLY3(4, apply_transposed_butterflies_r4_3d, OPK_TRANS)

/* ---- ordered transforms (fft.h:121-185) ----------------------------------- */
void fft(double* restrict re, double* restrict im, int n, int r, double** tr, double** ti, const int* rho) {
  (void)tr, (void)ti;
  D1(n);
  dropin(OPK_FFT, 1, dims_, r, re, im, rho, NULL, NULL);
}
void fft2(double* restrict re, double* restrict im, int n, int r, double** tr, double** ti, const int* rho) {
  (void)tr, (void)ti;
  D1(n);
  dropin(OPK_FFT2, 1, dims_, r, re, im, rho, NULL, NULL);
}
void ufft(double* restrict re, double* restrict im, int n, int r, double** tr, double** ti) {
  (void)tr, (void)ti;
  D1(n);
  dropin(OPK_UFFT, 1, dims_, r, re, im, NULL, NULL, NULL);
}
void ifft(double* restrict re, double* restrict im, int n, int r, double** tr, double** ti, const int* rho) {
  (void)tr, (void)ti;
  D1(n);
  dropin(OPK_IFFT, 1, dims_, r, re, im, rho, NULL, NULL);
}
void uifft(double* restrict re, double* restrict im, int n, int r, double** tr, double** ti) {
  (void)tr, (void)ti;
  D1(n);
  dropin(OPK_UIFFT, 1, dims_, r, re, im, NULL, NULL, NULL);
}

#define T2ARGS double** tr1, double** ti1, double** tr2, double** ti2
#define T2VOID (void)tr1, (void)ti1, (void)tr2, (void)ti2
void fft_2d(double* restrict re, double* restrict im, int n1, int n2, int r, T2ARGS, const int* rho) {
  T2VOID;
  D2(n1, n2);
  dropin(OPK_FFT, 2, dims_, r, re, im, rho, NULL, NULL);
}
void fft2_2d(double* restrict re, double* restrict im, int n1, int n2, int r, T2ARGS, const int* rho) {
  T2VOID;
  D2(n1, n2);
  dropin(OPK_FFT2, 2, dims_, r, re, im, rho, NULL, NULL);
}
void ufft_2d(double* restrict re, double* restrict im, int n1, int n2, int r, T2ARGS) {
  T2VOID;
  D2(n1, n2);
  dropin(OPK_UFFT, 2, dims_, r, re, im, NULL, NULL, NULL);
}
void ifft_2d(double* restrict re, double* restrict im, int n1, int n2, int r, T2ARGS, const int* rho) {
  T2VOID;
  D2(n1, n2);
  dropin(OPK_IFFT, 2, dims_, r, re, im, rho, NULL, NULL);
}
void uifft_2d(double* restrict re, double* restrict im, int n1, int n2, int r, T2ARGS) {
  T2VOID;
  D2(n1, n2);
  dropin(OPK_UIFFT, 2, dims_, r, re, im, NULL, NULL, NULL);
}

#define T3ARGS double** tr1, double** ti1, double** tr2, double** ti2, double** tr3, double** ti3
#define T3VOID (void)tr1, (void)ti1, (void)tr2, (void)ti2, (void)tr3, (void)ti3
void fft_3d(double* restrict re, double* restrict im, int n1, int n2, int n3, int r, T3ARGS, const int* rho) {
  T3VOID;
  D3(n1, n2, n3);
  dropin(OPK_FFT, 3, dims_, r, re, im, rho, NULL, NULL);
}
void fft2_3d(double* restrict re, double* restrict im, int n1, int n2, int n3, int r, T3ARGS, const int* rho) {
  T3VOID;
  D3(n1, n2, n3);
  dropin(OPK_FFT2, 3, dims_, r, re, im, rho, NULL, NULL);
}
void ufft_3d(double* restrict re, double* restrict im, int n1, int n2, int n3, int r, T3ARGS) {
  T3VOID;
  D3(n1, n2, n3);
  dropin(OPK_UFFT, 3, dims_, r, re, im, NULL, NULL, NULL);
}
void ifft_3d(double* restrict re, double* restrict im, int n1, int n2, int n3, int r, T3ARGS, const int* rho) {
  T3VOID;
  D3(n1, n2, n3);
  dropin(OPK_IFFT, 3, dims_, r, re, im, rho, NULL, NULL);
}
void uifft_3d(double* restrict re, double* restrict im, int n1, int n2, int n3, int r, T3ARGS) {
  T3VOID;
  D3(n1, n2, n3);
  dropin(OPK_UIFFT, 3, dims_, r, re, im, NULL, NULL, NULL);
}

/* ---- convolutions (fft.h:188-221) ----------------------------------------- */
void fft_based_convolution(double* restrict re, double* restrict im, int n, int r, double** tr, double** ti,
                           const int* rho, double* restrict g_re, double* restrict g_im) {
  (void)tr, (void)ti;
  D1(n);
  dropin(OPK_FB, 1, dims_, r, re, im, rho, g_re, g_im);
}
void permutation_avoiding_convolution(double* restrict re, double* restrict im, int n, int r, double** tr,
                                      double** ti, double* restrict g_re, double* restrict g_im) {
  (void)tr, (void)ti;
  D1(n);
  dropin(OPK_PA, 1, dims_, r, re, im, NULL, g_re, g_im);
}
void fft_based_convolution_2d(double* restrict re, double* restrict im, int n1, int n2, int r, T2ARGS,
                              const int* rho, double* restrict g_re, double* restrict g_im) {
  T2VOID;
  D2(n1, n2);
  dropin(OPK_FB, 2, dims_, r, re, im, rho, g_re, g_im);
}
void permutation_avoiding_convolution_2d(double* restrict re, double* restrict im, int n1, int n2, int r, T2ARGS,
                                         double* restrict g_re, double* restrict g_im) {
  T2VOID;
  D2(n1, n2);
  dropin(OPK_PA, 2, dims_, r, re, im, NULL, g_re, g_im);
}
void fft_based_convolution_3d(double* restrict re, double* restrict im, int n1, int n2, int n3, int r, T3ARGS,
                              const int* rho, double* restrict g_re, double* restrict g_im) {
  T3VOID;
  D3(n1, n2, n3);
  dropin(OPK_FB, 3, dims_, r, re, im, rho, g_re, g_im);
}
void permutation_avoiding_convolution_3d(double* restrict re, double* restrict im, int n1, int n2, int n3, int r,
                                         T3ARGS, double* restrict g_re, double* restrict g_im) {
  T3VOID;
  D3(n1, n2, n3);
  dropin(OPK_PA, 3, dims_, r, re, im, NULL, g_re, g_im);
}
