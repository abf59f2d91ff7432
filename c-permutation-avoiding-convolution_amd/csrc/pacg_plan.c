/*
 * Plans: schedules, pass lists, device twiddle tables and the launch
 * sequences of the permutation-avoiding convolution (host side, C11).
 *
 * A plan fixes the array shape, one stage schedule per axis and the working
 * precision.  Each axis' stages are cut into passes (runs of stages done in
 * one read-modify-write of the array, see pacg_pass.hpp); the launch
 * sequences below chain those passes:
 *
 *   A^T     (reference apply_transposed_butterflies_*): axes nd-1 .. 0, each
 *           axis top pass first (DIF order, butterflies.c:74, :226, :520);
 *   conj A / A: the reverse sequence, bottom pass first (butterflies.c:45, :16);
 *   convolve  A^T on x up to the last pass, which is fused with the spectral
 *           product and the first conj(A) pass, then the rest of conj(A)
 *           (convolution.c:31-72; 2D/3D: convolution_2d.c:35-70, _3d.c:36-72).
 * The fused pass works on positions that the forward and inverse stages
 * share, which is where the digit-reversal permutation cancels.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pacg.h"
#include "pacg_internal.h"
#include "pacg_host.h"

/* ---- errors --------------------------------------------------------------- */
static __thread char g_err[512];

const char* pacg_last_error(void) { return g_err; }
const char* pacg_version(void) { return "pacg 0.1 (gfx950)"; }

int pacg_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
void pacg_clear_error(void) { g_err[0] = 0; }

static int hip_check(int rc, const char* what) {
  if (rc != 0) return pacg_fail(PACG_ERR_HIP, "%s: HIP error %d (%s)", what, rc, pacg_hip_error_string(rc));
  return PACG_OK;
}
#define HIPCHK(call, what)                   \
  do {                                       \
    int rc_ = hip_check((call), (what));     \
    if (rc_ != PACG_OK) return rc_;          \
  } while (0)

/* ---- small integer helpers ------------------------------------------------ */
static int ilog2(long long v) {
  int l = 0;
  while ((1LL << l) < v) ++l;
  return ((1LL << l) == v) ? l : -1;
}
static int lg_radix(int r) { return r == 2 ? 1 : r == 4 ? 2 : r == 8 ? 3 : -1; }

/* ---- plan structures ------------------------------------------------------ */
#define MAX_PASS 16

typedef struct {
  int q_lo, q_hi, P, L;
  uint32_t code;
  int layout;
  int mfma; /* FP32 kernels of this pass come from the MFMA table */
  const pacg_variant_t* v[2][4][2]; /* [prec][kind][spec prec] */
  void* twp[2];
} pass_t;

typedef struct {
  int N, t;
  int radix[PACG_MAX_STAGES];
  int npass;
  pass_t pass[MAX_PASS]; /* pass[0] holds stage 0 */
  void* tw_hi[2];
  void* tw_lo[2];
  int lo_bits;
  long long inner, outer;
} axis_t;

typedef struct {
  int axis, pass;
} step_t;

struct pacg_plan {
  pacg_desc d;
  int nd;
  long long n;
  axis_t ax[3];
  int fused_axis;
  int nsteps;
  step_t steps[3 * MAX_PASS]; /* A^T order */
  void* stream;
  int own_stream;
  void* spec_re;
  void* spec_im;
  int has_spec;
  void* scr_re;
  void* scr_im;
  void* graph;
  int stream_x; /* the first A^T launch of a convolution loads x non-temporally (set by the autotuner) */
  int precs[2]; /* which precisions have tables */
  long long batch; /* length of a batch axis 0 (1 when none) */
};

static size_t esize(int prec) { return prec == PACG_F64 ? 8u : 4u; }
static int pidx(const pacg_plan* pl);
static const step_t* steps_for(const pacg_plan* pl, int prec, int* n);

static const pacg_variant_t* find_mfma(int spec_prec, int kind, int layout, int ns, uint32_t code) {
  for (int i = 0; i < pacg_num_mfma_variants; ++i) {
    const pacg_variant_t* v = &pacg_mfma_variants[i];
    if (v->rank == 0 && v->kind == kind && v->layout == layout && v->nstages == ns && v->stage_code == code &&
        (kind != PACG_K_FUSED || v->spec_prec == spec_prec))
      return v;
  }
  return NULL;
}

static const pacg_variant_t* find_variant(int prec, int spec_prec, int kind, int layout, int ns, uint32_t code) {
  for (int t = 0; t < pacg_num_variant_tables; ++t) {
    const pacg_variant_t* tab = pacg_variant_tables[t];
    for (int i = 0; i < pacg_variant_counts[t]; ++i) {
      const pacg_variant_t* v = &tab[i];
      if (v->rank == 0 && v->prec == prec && v->kind == kind && v->layout == layout && v->nstages == ns &&
          v->stage_code == code && (kind != PACG_K_FUSED || v->spec_prec == spec_prec))
        return v;
    }
  }
  return NULL;
}

static int same_pass(const pacg_variant_t* a, const pacg_variant_t* b);

/* Start a pass on a narrower tile of the same pass when the default tile
 * leaves CUs idle: of `lines` lines, the compiled tile with the most lines
 * per workgroup that still gives FILL_BLOCKS workgroups, else the one that
 * gives the most. */
enum { FILL_BLOCKS = 256 };
static const pacg_variant_t* fill_tile(const pacg_variant_t* v, long long lines) {
  if (!v || (lines + v->lines - 1) / v->lines >= FILL_BLOCKS) return v;
  const pacg_variant_t* best = v;
  for (int t = 0; t < pacg_num_variant_tables; ++t)
    for (int i = 0; i < pacg_variant_counts[t]; ++i) {
      const pacg_variant_t* c = &pacg_variant_tables[t][i];
      if (!same_pass(c, v)) continue;
      const long long bc = (lines + c->lines - 1) / c->lines, bb = (lines + best->lines - 1) / best->lines;
      if (bb >= FILL_BLOCKS ? (bc >= FILL_BLOCKS && c->lines > best->lines) : bc > bb) best = c;
    }
  return best;
}

static uint32_t run_code(const int* radix, int lo, int hi) {
  uint32_t c = 0;
  for (int q = lo; q <= hi; ++q) c |= (uint32_t)lg_radix(radix[q]) << (2 * (q - lo));
  return c;
}

int pacg_desc_init(pacg_desc* d, int nd, const int* dims, int r, int precision) {
  if (!d || !dims || nd < 1 || nd > 3) return pacg_fail(PACG_ERR_ARG, "pacg_desc_init: bad nd or NULL");
  if (lg_radix(r) < 0) return pacg_fail(PACG_ERR_ARG, "pacg_desc_init: radix %d not in {2,4,8}", r);
  memset(d, 0, sizeof *d);
  d->nd = nd;
  d->precision = precision;
  d->spectrum_precision = PACG_PREC_F64;
  d->fused_axis = -1;
  for (int a = 0; a < nd; ++a) {
    d->dims[a] = dims[a];
    long long m = 1;
    int t = 0;
    while (m < dims[a] && t < PACG_MAX_STAGES) {
      m *= r;
      d->radix[a][t++] = r;
    }
    if (m != dims[a])
      return pacg_fail(PACG_ERR_SIZE, "pacg_desc_init: axis %d length %d is not a power of %d", a, dims[a], r);
    d->nstages[a] = t;
  }
  return PACG_OK;
}

static int default_schedule(int N, int* radix) {
  int lg = ilog2(N), t = 0;
  if (lg < 0) return -1;
  for (int i = 0; i + 1 < lg; i += 2) radix[t++] = 4;
  if (lg & 1) radix[t++] = 2;
  return t;
}

/* Cut one axis' stages into passes, bottom up, taking the longest run with a
 * compiled kernel for the layout the pass will have (up to 512 points, or up
 * to 2048 where that leaves at least 1024 lines: on contiguous lines, the
 * bottom pass of the last axis, and on strided FP64 lines of the runs
 * compiled for them).  Single stages always have one. */
static int split_axis(axis_t* ax, int need_f32, int mfma) {
  int q = 0;
  ax->npass = 0;
  while (q < ax->t) {
    int best = 0, is_mfma = 0;
    if (mfma) {
      /* the longest run with a radix-8 stage that has an MFMA pass kernel (its
       * radix-8 stages on the matrix cores, the others on the VALU) */
      for (int len = ax->t - q; len >= 1 && !best; --len) {
        int has8 = 0;
        long long P = 1;
        for (int i = q; i < q + len; ++i) {
          has8 |= ax->radix[i] == 8;
          P *= ax->radix[i];
        }
        if (!has8 || P > 512) continue;
        uint32_t code = run_code(ax->radix, q, q + len - 1);
        if (!find_mfma(PACG_F32, PACG_K_DIF, PACG_L_STRIDED, len, code)) continue;
        if (!find_variant(PACG_F64, PACG_F64, PACG_K_DIF, PACG_L_STRIDED, len, code)) continue;
        best = len;
        is_mfma = 1;
      }
    }
    /* this pass' lines are contiguous; passes longer than 512 points only where they
     * still leave 1024 or more lines to spread over the CUs (and, strided, only in
     * FP64: need_f32 finds no FP32 kernel for them) */
    const int contig = q == 0 && ax->inner == 1;
    const long long npts = ax->outer * (long long)ax->N * ax->inner;
    for (int len = ax->t - q; len >= 1 && !best; --len) {
      if (mfma) {
        int has8 = 0;
        for (int i = q; i < q + len; ++i) has8 |= ax->radix[i] == 8;
        if (has8) continue;
      }
      if (len > 12) continue;
      long long P = 1;
      for (int i = q; i < q + len; ++i) P *= ax->radix[i];
      if (P > 2048 || (P > 512 && npts / P < 1024)) continue;
      uint32_t code = run_code(ax->radix, q, q + len - 1);
      int lay = PACG_L_STRIDED;
      if (!find_variant(PACG_F64, PACG_F64, PACG_K_DIF, lay, len, code)) {
        if (!contig) continue;
        lay = PACG_L_CONTIG;
        if (!find_variant(PACG_F64, PACG_F64, PACG_K_DIF, lay, len, code)) continue;
      }
      if (need_f32 && !find_variant(PACG_F32, PACG_F32, PACG_K_DIF, lay, len, code)) continue;
      best = len;
      break;
    }
    if (!best) return pacg_fail(PACG_ERR_NOKERNEL, "no pass kernel for stage %d (radix %d)", q, ax->radix[q]);
    if (ax->npass >= MAX_PASS) return pacg_fail(PACG_ERR_NOKERNEL, "too many passes on one axis");
    pass_t* ps = &ax->pass[ax->npass++];
    memset(ps, 0, sizeof *ps);
    ps->q_lo = q;
    ps->q_hi = q + best - 1;
    ps->P = 1;
    for (int i = ps->q_lo; i <= ps->q_hi; ++i) ps->P *= ax->radix[i];
    ps->L = 1;
    for (int i = 0; i < ps->q_lo; ++i) ps->L *= ax->radix[i];
    ps->code = run_code(ax->radix, ps->q_lo, ps->q_hi);
    ps->mfma = is_mfma;
    q += best;
  }
  return PACG_OK;
}

static int bind_variants(pacg_plan* pl) {
  for (int a = 0; a < pl->nd; ++a) {
    axis_t* ax = &pl->ax[a];
    for (int p = 0; p < ax->npass; ++p) {
      pass_t* ps = &ax->pass[p];
      long long S = (long long)ps->L * ax->inner;
      ps->layout = (S == 1) ? PACG_L_CONTIG : PACG_L_STRIDED;
      int ns = ps->q_hi - ps->q_lo + 1;
      long long P = 1;
      for (int q = ps->q_lo; q <= ps->q_hi; ++q) P *= ax->radix[q];
      for (int pr = 0; pr < 2; ++pr) {
        if (!pl->precs[pr]) continue;
        for (int k = 0; k < 4; ++k)
          for (int sp = 0; sp < 2; ++sp) {
            const pacg_variant_t* v = NULL;
            if (ps->mfma && pr == PACG_F32) {
              v = find_mfma(sp, k, ps->layout, ns, ps->code);
            } else {
              v = find_variant(pr, sp, k, ps->layout, ns, ps->code);
              if (!v && ps->layout == PACG_L_CONTIG) v = find_variant(pr, sp, k, PACG_L_STRIDED, ns, ps->code);
              v = fill_tile(v, pl->n / P);
            }
            ps->v[pr][k][sp] = v;
          }
        if (!ps->v[pr][PACG_K_DIF][pr] || !ps->v[pr][PACG_K_DITC][pr] || !ps->v[pr][PACG_K_DIT][pr])
          return pacg_fail(PACG_ERR_NOKERNEL, "missing pass kernel (axis %d pass %d)", a, p);
      }
    }
  }
  return PACG_OK;
}

static int build_tables(pacg_plan* pl) {
  for (int pr = 0; pr < 2; ++pr) {
    if (!pl->precs[pr]) continue;
    size_t es = 2 * esize(pr);
    for (int a = 0; a < pl->nd; ++a) {
      axis_t* ax = &pl->ax[a];
      if (ax->t == 0) continue;
      uint32_t lgN = (uint32_t)ilog2(ax->N);
      ax->lo_bits = (int)((lgN + 1) / 2);
      uint32_t nlo = 1u << ax->lo_bits, nhi = (uint32_t)ax->N >> ax->lo_bits;
      if (nhi == 0) nhi = 1;
      if (pacg_dev_malloc(&ax->tw_hi[pr], nhi * es) || pacg_dev_malloc(&ax->tw_lo[pr], nlo * es))
        return pacg_fail(PACG_ERR_NOMEM, "device allocation of twiddle tables failed");
      HIPCHK(pacg_gen_twiddles(ax->tw_hi[pr], pr, (uint32_t)ax->N, nhi, nlo, pl->stream), "twiddle generation");
      HIPCHK(pacg_gen_twiddles(ax->tw_lo[pr], pr, (uint32_t)ax->N, nlo, 1, pl->stream), "twiddle generation");
      for (int p = 0; p < ax->npass; ++p) {
        pass_t* ps = &ax->pass[p];
        if (pacg_dev_malloc(&ps->twp[pr], (size_t)ps->P * es))
          return pacg_fail(PACG_ERR_NOMEM, "device allocation of twiddle tables failed");
        HIPCHK(pacg_gen_twiddles(ps->twp[pr], pr, (uint32_t)ps->P, (uint32_t)ps->P, 1, pl->stream),
               "twiddle generation");
      }
    }
  }
  HIPCHK(pacg_stream_sync(pl->stream), "plan setup");
  return PACG_OK;
}

int pacg_plan_destroy(pacg_plan* pl) {
  if (!pl) return PACG_OK;
  if (pl->stream) pacg_stream_sync(pl->stream);
  for (int pr = 0; pr < 2; ++pr)
    for (int a = 0; a < pl->nd; ++a) {
      pacg_dev_free(pl->ax[a].tw_hi[pr]);
      pacg_dev_free(pl->ax[a].tw_lo[pr]);
      for (int p = 0; p < pl->ax[a].npass; ++p) pacg_dev_free(pl->ax[a].pass[p].twp[pr]);
    }
  pacg_dev_free(pl->spec_re);
  pacg_dev_free(pl->spec_im);
  pacg_dev_free(pl->scr_re);
  pacg_dev_free(pl->scr_im);
  if (pl->graph) pacg_graph_destroy(pl->graph);
  if (pl->own_stream && pl->stream) pacg_stream_destroy(pl->stream);
  free(pl);
  return PACG_OK;
}

int pacg_plan_create(pacg_plan** out, const pacg_desc* d) {
  if (!out || !d) return pacg_fail(PACG_ERR_ARG, "pacg_plan_create: NULL argument");
  *out = NULL;
  if (d->nd < 1 || d->nd > 3) return pacg_fail(PACG_ERR_ARG, "pacg_plan_create: nd = %d", d->nd);
  if (d->precision != PACG_PREC_F64 && d->precision != PACG_PREC_F32)
    return pacg_fail(PACG_ERR_ARG, "pacg_plan_create: precision %d", d->precision);
  pacg_plan* pl = (pacg_plan*)calloc(1, sizeof *pl);
  if (!pl) return pacg_fail(PACG_ERR_NOMEM, "pacg_plan_create: out of host memory");
  pl->d = *d;
  pl->nd = d->nd;
  pl->n = 1;
  for (int a = 0; a < d->nd; ++a) {
    if (d->nstages[a] == -1 && a == 0 && d->dims[a] >= 1) {
      pl->n *= d->dims[a];
      continue;
    }
    if (d->dims[a] < 1 || ilog2(d->dims[a]) < 0) {
      free(pl);
      return pacg_fail(PACG_ERR_SIZE, "axis %d length %d is not a power of two", a, d->dims[a]);
    }
    pl->n *= d->dims[a];
  }
  if (pl->n >= (1LL << 31)) {
    free(pl);
    return pacg_fail(PACG_ERR_SIZE, "total size %lld >= 2^31", pl->n);
  }
  pl->batch = 1;
  pl->precs[PACG_F64] = 1;
  pl->precs[PACG_F32] = d->precision == PACG_PREC_F32;
  int rc = PACG_OK;
  for (int a = 0; a < d->nd; ++a) {
    axis_t* ax = &pl->ax[a];
    ax->N = d->dims[a];
    ax->inner = 1;
    for (int b = a + 1; b < d->nd; ++b) ax->inner *= d->dims[b];
    ax->outer = pl->n / ((long long)ax->N * ax->inner);
    if (d->nstages[a] == -1) {
      if (a != 0) {
        rc = pacg_fail(PACG_ERR_ARG, "only axis 0 can be a batch axis");
        goto fail;
      }
      ax->t = 0;
      pl->batch = ax->N;
      continue;
    } else if (d->nstages[a] == 0) {
      ax->t = default_schedule(ax->N, ax->radix);
    } else {
      if (d->nstages[a] < 0 || d->nstages[a] > PACG_MAX_STAGES) {
        rc = pacg_fail(PACG_ERR_ARG, "axis %d: %d stages", a, d->nstages[a]);
        goto fail;
      }
      long long prod = 1;
      ax->t = d->nstages[a];
      for (int q = 0; q < ax->t; ++q) {
        if (lg_radix(d->radix[a][q]) < 0) {
          rc = pacg_fail(PACG_ERR_ARG, "axis %d stage %d: radix %d not in {2,4,8}", a, q, d->radix[a][q]);
          goto fail;
        }
        ax->radix[q] = d->radix[a][q];
        prod *= ax->radix[q];
      }
      if (prod != ax->N) {
        rc = pacg_fail(PACG_ERR_SIZE, "axis %d: schedule product %lld != length %d", a, prod, ax->N);
        goto fail;
      }
    }
    if (ax->N == 1) ax->t = 0;
    if ((rc = split_axis(ax, pl->precs[PACG_F32], d->mfma && d->precision == PACG_PREC_F32)) != PACG_OK) goto fail;
  }
  /* fused axis: requested one if it has stages, else the first that has */
  pl->fused_axis = -1;
  if (d->fused_axis >= 0 && d->fused_axis < d->nd && pl->ax[d->fused_axis].t > 0) pl->fused_axis = d->fused_axis;
  for (int a = 0; a < d->nd && pl->fused_axis < 0; ++a)
    if (pl->ax[a].t > 0) pl->fused_axis = a;
  /* ... unless, not requested, the fused pass there would run strided lines of
   * more than 512 points: such tiles hold 2-4 lines (32 KiB per FP64 line of
   * 2048 points), and with data and spectrum the fused form gets only the 2-line
   * tile (2048^2 FP64: 85 us against 48 us for the fused pass on contiguous
   * lines).  The product then moves to the last axis with stages. */
  if (d->fused_axis < 0 && pl->fused_axis >= 0) {
    const axis_t* fx = &pl->ax[pl->fused_axis];
    if (fx->npass > 0 && fx->pass[0].P > 512 && fx->inner > 1)
      for (int a = d->nd - 1; a > pl->fused_axis; --a)
        if (pl->ax[a].t > 0) {
          pl->fused_axis = a;
          break;
        }
  }
  /* A^T step order: non-fused axes innermost first, then the fused axis; top pass first */
  pl->nsteps = 0;
  for (int a = d->nd - 1; a >= 0; --a) {
    if (a == pl->fused_axis) continue;
    for (int p = pl->ax[a].npass - 1; p >= 0; --p) pl->steps[pl->nsteps++] = (step_t){a, p};
  }
  if (pl->fused_axis >= 0)
    for (int p = pl->ax[pl->fused_axis].npass - 1; p >= 0; --p)
      pl->steps[pl->nsteps++] = (step_t){pl->fused_axis, p};
  if ((rc = bind_variants(pl)) != PACG_OK) goto fail;
  if (pacg_stream_create(&pl->stream)) {
    rc = pacg_fail(PACG_ERR_HIP, "stream creation failed");
    goto fail;
  }
  pl->own_stream = 1;
  if ((rc = build_tables(pl)) != PACG_OK) goto fail;
  {
    int sp = d->spectrum_precision == PACG_PREC_F32 ? PACG_F32 : PACG_F64;
    if (d->precision == PACG_PREC_F64) sp = PACG_F64;
    pl->d.spectrum_precision = sp;
  }
  /* set-up of the allocation-free entry point on this device: resolve the
   * address of its fixed scratch now (a failure here is reported by that
   * entry point, which resolves it again) */
  {
    void* scr = NULL;
    (void)pacg_noalloc_scratch(&scr);
  }
  *out = pl;
  pacg_clear_error();
  return PACG_OK;
fail:
  pacg_plan_destroy(pl);
  return rc;
}

int pacg_plan_set_stream(pacg_plan* pl, void* stream) {
  if (!pl) return pacg_fail(PACG_ERR_ARG, "NULL plan");
  if (pl->own_stream && pl->stream) {
    pacg_stream_sync(pl->stream);
    pacg_stream_destroy(pl->stream);
  }
  pl->own_stream = 0;
  pl->stream = stream;
  return PACG_OK;
}
void* pacg_plan_stream(const pacg_plan* pl) { return pl ? pl->stream : NULL; }
long long pacg_plan_size(const pacg_plan* pl) { return pl ? pl->n : 0; }
int pacg_hbm_probe(int mode, void* dst, const void* src, uint64_t bytes, int reps, void* stream, float* ms) {
  if (mode < 0 || mode > 2) return pacg_fail(PACG_ERR_ARG, "pacg_hbm_probe: mode %d not in 0..2", mode);
  if (bytes == 0 || bytes % 16 != 0)
    return pacg_fail(PACG_ERR_ARG, "pacg_hbm_probe: %llu bytes is not a positive multiple of 16",
                     (unsigned long long)bytes);
  if (reps < 1 || !ms || (mode != 1 && !src) || (mode != 0 && !dst))
    return pacg_fail(PACG_ERR_ARG, "pacg_hbm_probe: bad argument");
  void *e0 = NULL, *e1 = NULL;
  int rc = PACG_OK;
  if (pacg_event_create(&e0) || pacg_event_create(&e1)) {
    rc = pacg_fail(PACG_ERR_HIP, "pacg_hbm_probe: event creation failed");
  } else if (pacg_bw_probe(mode, dst, src, bytes, stream)) { /* warm-up */
    rc = pacg_fail(PACG_ERR_HIP, "pacg_hbm_probe: launch failed");
  } else {
    int e = pacg_event_record(e0, stream);
    for (int r = 0; r < reps && !e; ++r) e |= pacg_bw_probe(mode, dst, src, bytes, stream);
    e |= pacg_event_record(e1, stream);
    e |= pacg_event_sync(e1);
    if (!e) e |= pacg_event_elapsed_ms(ms, e0, e1);
    if (e) rc = pacg_fail(PACG_ERR_HIP, "pacg_hbm_probe: timing failed (%d)", e);
    else *ms /= (float)reps;
  }
  if (e0) pacg_event_destroy(e0);
  if (e1) pacg_event_destroy(e1);
  if (rc == PACG_OK) pacg_clear_error();
  return rc;
}

int pacg_sync(pacg_plan* pl) {
  if (!pl) return pacg_fail(PACG_ERR_ARG, "NULL plan");
  HIPCHK(pacg_stream_sync(pl->stream), "stream synchronize");
  return PACG_OK;
}

int pacg_plan_rho(const pacg_plan* pl, int* rho) {
  if (!pl || !rho) return pacg_fail(PACG_ERR_ARG, "NULL argument");
  for (long long idx = 0; idx < pl->n; ++idx) {
    long long rem = idx, acc = 0, mul = 1;
    for (int a = pl->nd - 1; a >= 0; --a) {
      const axis_t* ax = &pl->ax[a];
      long long i = rem % ax->N;
      rem /= ax->N;
      acc += (ax->t ? pacg_rev_index_mixed(i, ax->N, ax->radix, ax->t) : i) * mul;
      mul *= ax->N;
    }
    rho[idx] = (int)acc;
  }
  return PACG_OK;
}

int pacg_plan_describe(const pacg_plan* pl, char* buf, int buflen) {
  if (!pl || !buf || buflen < 1) return pacg_fail(PACG_ERR_ARG, "NULL argument");
  int off = 0, ns;
  const step_t* S = steps_for(pl, pidx(pl), &ns);
  buf[0] = 0;
  for (int s = 0; s < ns && off < buflen; ++s) {
    const axis_t* ax = &pl->ax[S[s].axis];
    const pass_t* ps = &ax->pass[S[s].pass];
    const pacg_variant_t* v = ps->v[pl->d.precision == PACG_PREC_F32][PACG_K_DIF][pl->d.precision == PACG_PREC_F32];
    off += snprintf(buf + off, (size_t)(buflen - off), "%saxis%d[q%d..%d P=%d L=%d %s]", s ? " " : "", S[s].axis,
                    ps->q_lo, ps->q_hi, ps->P, ps->L, v ? v->name : "?");
  }
  if (pl->stream_x && off < buflen) snprintf(buf + off, (size_t)(buflen - off), " x:nt");
  return ns > 0 ? 2 * ns - 1 : 1;
}

/* ---- launching ------------------------------------------------------------ */
static int launch_step(pacg_plan* pl, step_t st, int prec, int kind, int sprec, const void* ir, const void* ii,
                       void* orr, void* oi, const void* sr, const void* si, double scale, double out_scale, int in_nt) {
  axis_t* ax = &pl->ax[st.axis];
  pass_t* ps = &ax->pass[st.pass];
  const pacg_variant_t* v = ps->v[prec][kind][sprec];
  if (!v) return pacg_fail(PACG_ERR_NOKERNEL, "no kernel for kind %d on axis %d pass %d", kind, st.axis, st.pass);
  pacg_pass_args_t A;
  memset(&A, 0, sizeof A);
  long long S = (long long)ps->L * ax->inner;
  A.in_re = ir;
  A.in_im = ii;
  A.out_re = orr;
  A.out_im = oi;
  A.spec_re = sr;
  A.spec_im = si;
  A.tw_p = ps->twp[prec];
  A.tw_hi = ax->tw_hi[prec];
  A.tw_lo = ax->tw_lo[prec];
  A.log2_S = v->layout == PACG_L_CONTIG ? 0u : (uint32_t)ilog2(S);
  A.n_o = (uint32_t)(ax->outer * (ax->N / ((long long)ps->P * ps->L)));
  if (v->layout == PACG_L_CONTIG) A.n_o = (uint32_t)(pl->n / ps->P);
  A.log2_inner = (uint32_t)ilog2(ax->inner);
  A.ext = ps->L > 1;
  A.ext_shift = (uint32_t)ilog2(ax->N / ((long long)ps->P * ps->L));
  A.lo_bits = (uint32_t)ax->lo_bits;
  A.scale = scale;
  A.out_scale = out_scale;
  A.vec16 = (((uintptr_t)ir | (uintptr_t)ii | (uintptr_t)orr | (uintptr_t)oi) & 15u) == 0;
  if (kind == PACG_K_FUSED && sprec == PACG_F64 && v->layout == PACG_L_STRIDED) {
    /* When the array and the spectrum together exceed the 256 MiB Infinity
     * Cache, an FP64-stored spectrum (read once per convolution) is streamed
     * past it, so that the array stays cached for the passes either side:
     * 256^3 FP32 with an FP64 spectrum (384 MiB) 0.290 -> 0.262 ms, 256^3
     * FP64 (512 MiB) 0.539 -> 0.481 ms.  Only on strided tiles, whose loads
     * read whole runs of lines per instruction: a contiguous tile reads each
     * element's points R apart, and without cache allocation every element
     * fetches the lines again (1D 2^24 FP64 fused pass 151 -> 216 us).  A
     * working set that fits (2048^2 FP64, 128 MiB) keeps the default loads,
     * as do FP32 spectra (non-temporal loads slowed the 256^3 FP32 pass). */
    const unsigned long long bytes = (unsigned long long)pl->n * 2ull * ((prec == PACG_F32 ? 4ull : 8ull) + 8ull);
    A.spec_nt = bytes > (256ull << 20);
  }
  A.in_nt = kind == PACG_K_DIF && in_nt;
  int rc = v->launch(&A, pl->stream);
  if (rc) return pacg_fail(PACG_ERR_HIP, "launch of %s failed: %s", v->name, pacg_hip_error_string(rc));
  return PACG_OK;
}

static int pidx(const pacg_plan* pl) { return pl->d.precision == PACG_PREC_F32 ? PACG_F32 : PACG_F64; }

/* launch steps (A^T order) */
static const step_t* steps_for(const pacg_plan* pl, int prec, int* n) {
  (void)prec;
  *n = pl->nsteps;
  return pl->steps;
}

/* A^T (kind DIF), A (DIT) or conj(A) (DITC) over all axes: in -> out. */
static int run_all(pacg_plan* pl, int prec, int kind, const void* ir, const void* ii, void* orr, void* oi,
                   double out_scale) {
  int ns;
  const step_t* S = steps_for(pl, prec, &ns);
  if (ns == 0) {
    HIPCHK(pacg_convert(ir, ii, prec, orr, oi, prec, (uint64_t)pl->n, out_scale, pl->stream), "copy");
    return PACG_OK;
  }
  for (int i = 0; i < ns; ++i) {
    step_t st = (kind == PACG_K_DIF) ? S[i] : S[ns - 1 - i];
    const void* a = i == 0 ? ir : orr;
    const void* b = i == 0 ? ii : oi;
    int rc = launch_step(pl, st, prec, kind, prec, a, b, orr, oi, NULL, NULL, 1.0, i == ns - 1 ? out_scale : 1.0, 0);
    if (rc) return rc;
  }
  return PACG_OK;
}

/* The launches of one convolution: A^T passes but the last, the fused pass,
 * then the conj(A) passes in reverse (convolution.c:35-71). */
static int conv_launch(const pacg_plan* pl, int i, step_t* st, int* kind) {
  int ns;
  const step_t* S = steps_for(pl, pidx(pl), &ns);
  if (i < ns - 1) {
    *st = S[i];
    *kind = PACG_K_DIF;
  } else if (i == ns - 1) {
    *st = S[ns - 1];
    *kind = PACG_K_FUSED;
  } else {
    *st = S[2 * ns - 2 - i];
    *kind = PACG_K_DITC;
  }
  return 2 * ns - 1;
}

/* launches of one convolution in the plan precision (0: no transform axes) */
static int conv_launches(const pacg_plan* pl) {
  int ns;
  steps_for(pl, pidx(pl), &ns);
  return ns ? 2 * ns - 1 : 0;
}

/* y = (1/n) conj(A)[spec .* (A^T x)]; events (if given) are recorded before
 * launch i and after the last one. */
static int run_conv_ev(pacg_plan* pl, int prec, const void* xr, const void* xi, void* yr, void* yi, const void* sr,
                       const void* si, int sprec, void** events) {
  double scale = (double)pl->batch / (double)pl->n;
  if (conv_launches(pl) == 0) {
    if (xr != yr) HIPCHK(pacg_convert(xr, xi, prec, yr, yi, prec, (uint64_t)pl->n, 1.0, pl->stream), "copy");
    HIPCHK(pacg_pointwise(yr, yi, sr, si, prec, sprec, (uint64_t)pl->n, scale, pl->stream), "pointwise");
    return PACG_OK;
  }
  int nl = conv_launches(pl);
  for (int i = 0; i < nl; ++i) {
    step_t st;
    int kind;
    conv_launch(pl, i, &st, &kind);
    if (events) HIPCHK(pacg_event_record(events[i], pl->stream), "event record");
    int rc = launch_step(pl, st, prec, kind, kind == PACG_K_FUSED ? sprec : prec, i == 0 ? xr : yr,
                         i == 0 ? xi : yi, yr, yi, kind == PACG_K_FUSED ? sr : NULL, kind == PACG_K_FUSED ? si : NULL,
                         kind == PACG_K_FUSED ? scale : 1.0, 1.0, i == 0 && pl->stream_x);
    if (rc) return rc;
  }
  if (events) HIPCHK(pacg_event_record(events[nl], pl->stream), "event record");
  return PACG_OK;
}

static int run_conv(pacg_plan* pl, int prec, const void* xr, const void* xi, void* yr, void* yi, const void* sr,
                    const void* si, int sprec) {
  return run_conv_ev(pl, prec, xr, xi, yr, yi, sr, si, sprec, NULL);
}

/* Autotuner: for every launch of one convolution, time each compiled tile
 * shape of its pass (same stages, kind, layout and precisions) on the given
 * buffers, out of place x -> y, and keep the fastest.  Returns the number of
 * launches whose kernel changed, or minus an error code. */
/* The kernel slot of launch i of a convolution and its current kernel (NULL
 * for an index out of range).  Any compiled kernel of the
 * same pass (same_pass: precision, kind, layout, stages, spectrum precision
 * and engine; any lines per workgroup) can take the slot. */
static const pacg_variant_t* launch_slot(pacg_plan* pl, int i, const pacg_variant_t*** slot) {
  step_t st;
  int kind;
  if (i < 0 || i >= conv_launches(pl)) return NULL;
  conv_launch(pl, i, &st, &kind);
  int prec = pidx(pl), vsp = kind == PACG_K_FUSED ? pl->d.spectrum_precision : prec;
  *slot = &pl->ax[st.axis].pass[st.pass].v[prec][kind][vsp];
  return **slot;
}

static int same_pass(const pacg_variant_t* a, const pacg_variant_t* b) {
  return a->prec == b->prec && a->kind == b->kind && a->layout == b->layout && a->nstages == b->nstages &&
         a->stage_code == b->stage_code && a->spec_prec == b->spec_prec && a->engine == b->engine;
}

int pacg_plan_launch_candidates(pacg_plan* pl, int launch, const char** names, int max_names) {
  if (!pl || max_names < 0 || (max_names > 0 && !names))
    return -pacg_fail(PACG_ERR_ARG, "pacg_plan_launch_candidates: bad argument");
  const pacg_variant_t** slot = NULL;
  const pacg_variant_t* cur = launch_slot(pl, launch, &slot);
  if (!cur) return -pacg_fail(PACG_ERR_ARG, "launch %d has no pass kernel to choose", launch);
  int n = 0;
  for (int t = 0; t <= pacg_num_variant_tables; ++t) {
    const pacg_variant_t* tab = t < pacg_num_variant_tables ? pacg_variant_tables[t] : pacg_mfma_variants;
    const int cnt = t < pacg_num_variant_tables ? pacg_variant_counts[t] : pacg_num_mfma_variants;
    for (int j = 0; j < cnt; ++j)
      if (same_pass(&tab[j], cur)) {
        if (n < max_names) names[n] = tab[j].name;
        ++n;
      }
  }
  return n;
}

int pacg_plan_pin_launch(pacg_plan* pl, int launch, const char* name) {
  if (!pl || !name) return pacg_fail(PACG_ERR_ARG, "pacg_plan_pin_launch: bad argument");
  const pacg_variant_t** slot = NULL;
  const pacg_variant_t* cur = launch_slot(pl, launch, &slot);
  if (!cur) return pacg_fail(PACG_ERR_ARG, "launch %d has no pass kernel to choose", launch);
  for (int t = 0; t <= pacg_num_variant_tables; ++t) {
    const pacg_variant_t* tab = t < pacg_num_variant_tables ? pacg_variant_tables[t] : pacg_mfma_variants;
    const int cnt = t < pacg_num_variant_tables ? pacg_variant_counts[t] : pacg_num_mfma_variants;
    for (int j = 0; j < cnt; ++j)
      if (same_pass(&tab[j], cur) && strcmp(tab[j].name, name) == 0) {
        *slot = &tab[j];
        if (pl->graph) {
          pacg_graph_destroy(pl->graph);
          pl->graph = NULL;
        }
        return PACG_OK;
      }
  }
  return pacg_fail(PACG_ERR_ARG, "'%s' is not a tile shape of launch %d (current: %s)", name, launch, cur->name);
}

int pacg_plan_stream_input(pacg_plan* pl, int on) {
  if (!pl) return pacg_fail(PACG_ERR_ARG, "pacg_plan_stream_input: NULL plan");
  pl->stream_x = on != 0;
  if (pl->graph) {
    pacg_graph_destroy(pl->graph);
    pl->graph = NULL;
  }
  return PACG_OK;
}

int pacg_plan_autotune(pacg_plan* pl, const void* xr, const void* xi, void* yr, void* yi, int reps) {
  if (!pl || !xr || !xi || !yr || !yi || reps < 1) return -pacg_fail(PACG_ERR_ARG, "pacg_plan_autotune: bad argument");
  if (!pl->has_spec) return -pacg_fail(PACG_ERR_STATE, "no kernel spectrum set");
  if (conv_launches(pl) == 0) return 0;
  int prec = pidx(pl), sp = pl->d.spectrum_precision, changed = 0;
  void *e0 = NULL, *e1 = NULL;
  if (pacg_event_create(&e0) || pacg_event_create(&e1)) return -pacg_fail(PACG_ERR_HIP, "event creation failed");
  int nl = conv_launches(pl), rc = PACG_OK;
  for (int i = 0; i < nl && !rc; ++i) {
    step_t st;
    int kind;
    conv_launch(pl, i, &st, &kind);
    pass_t* ps = &pl->ax[st.axis].pass[st.pass];
    int vsp = kind == PACG_K_FUSED ? sp : prec;
    const pacg_variant_t* cur = ps->v[prec][kind][vsp];
    if (!cur) continue;
    /* candidates: the generated pass kernels, then the MFMA pass kernels, of the same pass */
    enum { MAXC = 32 };
    const pacg_variant_t* cand[MAXC];
    float cms[MAXC];
    int nc = 0, def = -1;
    for (int t = 0; t <= pacg_num_variant_tables && nc < MAXC; ++t) {
      const pacg_variant_t* tab = t < pacg_num_variant_tables ? pacg_variant_tables[t] : pacg_mfma_variants;
      const int cnt = t < pacg_num_variant_tables ? pacg_variant_counts[t] : pacg_num_mfma_variants;
      for (int j = 0; j < cnt && nc < MAXC; ++j) {
        if (!same_pass(&tab[j], cur)) continue;
        if (&tab[j] == cur) def = nc;
        cms[nc] = 1e30f;
        cand[nc++] = &tab[j];
      }
    }
    /* two rounds over the candidates, each keeps its faster round: a drift of the
     * clock during tuning then penalises no candidate for its place in the list */
    for (int round = 0; round < 2 && !rc; ++round) {
      for (int c = 0; c < nc && !rc; ++c) {
        const pacg_variant_t* v = cand[c];
        ps->v[prec][kind][vsp] = v;
        const void* s_re = kind == PACG_K_FUSED ? pl->spec_re : NULL;
        const void* s_im = kind == PACG_K_FUSED ? pl->spec_im : NULL;
        double scale = kind == PACG_K_FUSED ? (double)pl->batch / (double)pl->n : 1.0;
        if ((rc = launch_step(pl, st, prec, kind, vsp, xr, xi, yr, yi, s_re, s_im, scale, 1.0, 0))) break;
        if (pacg_event_record(e0, pl->stream)) rc = pacg_fail(PACG_ERR_HIP, "event record failed");
        for (int r = 0; r < reps && !rc; ++r)
          rc = launch_step(pl, st, prec, kind, vsp, xr, xi, yr, yi, s_re, s_im, scale, 1.0, 0);
        if (rc) break;
        if (pacg_event_record(e1, pl->stream) || pacg_stream_sync(pl->stream)) {
          rc = pacg_fail(PACG_ERR_HIP, "autotune timing failed");
          break;
        }
        float ms = 0.0f;
        pacg_event_elapsed_ms(&ms, e0, e1);
        if (ms < cms[c]) cms[c] = ms;
      }
    }
    /* the fastest; the tile the launch had unless another is faster by more than 2 % */
    const pacg_variant_t* best = cur;
    if (nc > 0) {
      int b = 0;
      for (int c = 1; c < nc; ++c)
        if (cms[c] < cms[b]) b = c;
      if (def >= 0 && cms[def] <= 1.02f * cms[b]) b = def;
      best = cand[b];
    }
    ps->v[prec][kind][vsp] = best;
    if (best != cur) ++changed;
  }
  /* Load policy of the first launch, an A^T pass that reads the caller's x
   * (read once per convolution): default or non-temporal loads, by timing
   * whole convolutions both ways (at least 8 per timing, three rounds, each
   * policy keeps its fastest round); the policy the plan had stays unless the
   * other is faster by more than 1 % (the gain is 2-8 %; with 3 convolutions
   * per timing and a 2 % margin, noise decided).  A streamed x leaves the
   * Infinity Cache to y and the spectrum, which the later launches read:
   * 256^3 FP32 0.235 -> 0.230 ms, 256^3 FP64 0.490 -> 0.451 ms. */
  step_t st0;
  int k0 = -1;
  const pacg_variant_t** slot0 = NULL;
  const pacg_variant_t* v0 = (!rc && nl >= 2) ? launch_slot(pl, 0, &slot0) : NULL;
  if (v0) conv_launch(pl, 0, &st0, &k0);
  if (k0 == PACG_K_DIF && v0->engine == 0) { /* MFMA pass kernels have one load policy */
    float tb[2] = {1e30f, 1e30f};
    const int had = pl->stream_x, nconv = reps > 8 ? reps : 8;
    for (int round = 0; round < 3 && !rc; ++round)
      for (int f = 0; f < 2 && !rc; ++f) {
        pl->stream_x = f;
        if ((rc = run_conv(pl, prec, xr, xi, yr, yi, pl->spec_re, pl->spec_im, sp))) break;
        if (pacg_event_record(e0, pl->stream)) {
          rc = pacg_fail(PACG_ERR_HIP, "event record failed");
          break;
        }
        for (int r = 0; r < nconv && !rc; ++r) rc = run_conv(pl, prec, xr, xi, yr, yi, pl->spec_re, pl->spec_im, sp);
        if (rc) break;
        if (pacg_event_record(e1, pl->stream) || pacg_stream_sync(pl->stream)) {
          rc = pacg_fail(PACG_ERR_HIP, "autotune timing failed");
          break;
        }
        float ms = 0.0f;
        pacg_event_elapsed_ms(&ms, e0, e1);
        if (ms < tb[f]) tb[f] = ms;
      }
    pl->stream_x = rc ? had : had ? !(tb[0] < 0.99f * tb[1]) : tb[1] < 0.99f * tb[0];
    if (pl->stream_x != had) ++changed;
  }
  pacg_event_destroy(e0);
  pacg_event_destroy(e1);
  if (pl->graph) {
    pacg_graph_destroy(pl->graph);
    pl->graph = NULL;
  }
  return rc ? -rc : changed;
}

int pacg_profile_convolve(pacg_plan* pl, const void* xr, const void* xi, void* yr, void* yi, int reps,
                          float* launch_ms, const char** names, int max_launches) {
  if (!pl || !xr || !xi || !yr || !yi || reps < 1 || !launch_ms)
    return -pacg_fail(PACG_ERR_ARG, "pacg_profile_convolve: bad argument");
  if (!pl->has_spec) return -pacg_fail(PACG_ERR_STATE, "no kernel spectrum set");
  int nl = conv_launches(pl);
  if (nl == 0 || nl > max_launches) return nl;
  void* ev[3 * MAX_PASS * 2 + 2];
  for (int i = 0; i <= nl; ++i)
    if (pacg_event_create(&ev[i])) return -pacg_fail(PACG_ERR_HIP, "event creation failed");
  int prec = pidx(pl), sp = pl->d.spectrum_precision;
  for (int i = 0; i < nl; ++i) launch_ms[i] = 0.0f;
  int rc = PACG_OK;
  for (int r = 0; r < reps && !rc; ++r) {
    rc = run_conv_ev(pl, prec, xr, xi, yr, yi, pl->spec_re, pl->spec_im, sp, ev);
    if (rc) break;
    if (pacg_stream_sync(pl->stream)) {
      rc = pacg_fail(PACG_ERR_HIP, "stream synchronize failed");
      break;
    }
    for (int i = 0; i < nl; ++i) {
      float ms = 0.0f;
      pacg_event_elapsed_ms(&ms, ev[i], ev[i + 1]);
      launch_ms[i] += ms / (float)reps;
    }
  }
  if (names)
    for (int i = 0; i < nl; ++i) {
      step_t st;
      int kind;
      conv_launch(pl, i, &st, &kind);
      int vsp = kind == PACG_K_FUSED ? sp : prec;
      const pacg_variant_t* v = pl->ax[st.axis].pass[st.pass].v[prec][kind][vsp];
      names[i] = v ? v->name : "?";
    }
  for (int i = 0; i <= nl; ++i) pacg_event_destroy(ev[i]);
  return rc ? -rc : nl;
}

static int ensure_scratch(pacg_plan* pl) {
  if (pl->scr_re) return PACG_OK;
  size_t bytes = (size_t)pl->n * esize(pidx(pl));
  if (pacg_dev_malloc(&pl->scr_re, bytes) || pacg_dev_malloc(&pl->scr_im, bytes))
    return pacg_fail(PACG_ERR_NOMEM, "device allocation of %zu bytes failed", 2 * bytes);
  return PACG_OK;
}

static void perm_desc(const pacg_plan* pl, pacg_perm_desc_t* pd) {
  memset(pd, 0, sizeof *pd);
  pd->nd = pl->nd;
  for (int a = 0; a < pl->nd; ++a) {
    pd->dims[a] = (uint32_t)pl->ax[a].N;
    pd->nst[a] = pl->ax[a].t;
    for (int q = 0; q < pl->ax[a].t; ++q) pd->radix_log2[a][q] = (uint8_t)lg_radix(pl->ax[a].radix[q]);
  }
}

/* pacg_execute with the digit-reversal scratch (sr, si) of n elements each,
 * used by the ordered transforms (FFT, IFFT, FFT2) only */
static int execute_with(pacg_plan* pl, int op, const void* ir, const void* ii, void* orr, void* oi, void* sr,
                        void* si) {
  int prec = pidx(pl);
  double inv_n = (double)pl->batch / (double)pl->n;
  pacg_perm_desc_t pd;
  int rc;
  switch (op) {
    case PACG_OP_TRANS:
      return run_all(pl, prec, PACG_K_DIF, ir, ii, orr, oi, 1.0);
    case PACG_OP_CONJ:
      return run_all(pl, prec, PACG_K_DITC, ir, ii, orr, oi, 1.0);
    case PACG_OP_FWD:
      return run_all(pl, prec, PACG_K_DIT, ir, ii, orr, oi, 1.0);
    case PACG_OP_UIFFT:
      return run_all(pl, prec, PACG_K_DITC, ir, ii, orr, oi, inv_n);
    case PACG_OP_FFT:
    case PACG_OP_IFFT:
      perm_desc(pl, &pd);
      HIPCHK(pacg_permute(ir, ii, sr, si, prec, &pd, 0, pl->stream), "permute");
      return run_all(pl, prec, op == PACG_OP_FFT ? PACG_K_DIT : PACG_K_DITC, sr, si, orr, oi,
                     op == PACG_OP_FFT ? 1.0 : inv_n);
    case PACG_OP_FFT2:
      perm_desc(pl, &pd);
      if ((rc = run_all(pl, prec, PACG_K_DIF, ir, ii, sr, si, 1.0))) return rc;
      HIPCHK(pacg_permute(sr, si, orr, oi, prec, &pd, 1, pl->stream), "permute");
      return PACG_OK;
    case PACG_OP_PERM:
      perm_desc(pl, &pd);
      if (ir != orr || ii != oi) {
        HIPCHK(pacg_permute(ir, ii, orr, oi, prec, &pd, 0, pl->stream), "permute");
        return PACG_OK;
      }
      HIPCHK(pacg_permute(ir, ii, sr, si, prec, &pd, 0, pl->stream), "permute");
      HIPCHK(pacg_convert(sr, si, prec, orr, oi, prec, (uint64_t)pl->n, 1.0, pl->stream), "copy");
      return PACG_OK;
    default:
      return pacg_fail(PACG_ERR_ARG, "pacg_execute: unknown op %d", op);
  }
}

int pacg_execute(pacg_plan* pl, int op, const void* ir, const void* ii, void* orr, void* oi) {
  if (!pl || !ir || !ii || !orr || !oi) return pacg_fail(PACG_ERR_ARG, "pacg_execute: NULL argument");
  int rc;
  if ((op == PACG_OP_FFT || op == PACG_OP_IFFT || op == PACG_OP_FFT2 || (op == PACG_OP_PERM && ir == orr)) &&
      (rc = ensure_scratch(pl)))
    return rc;
  return execute_with(pl, op, ir, ii, orr, oi, pl->scr_re, pl->scr_im);
}

int pacg_set_kernel(pacg_plan* pl, const double* kr, const double* ki, int on_device) {
  if (!pl || !kr || !ki) return pacg_fail(PACG_ERR_ARG, "pacg_set_kernel: NULL argument");
  size_t n = (size_t)pl->n;
  int sp = pl->d.spectrum_precision;
  if (!pl->spec_re) {
    if (pacg_dev_malloc(&pl->spec_re, n * esize(sp)) || pacg_dev_malloc(&pl->spec_im, n * esize(sp)))
      return pacg_fail(PACG_ERR_NOMEM, "device allocation of the spectrum failed");
  }
  void *t_re = NULL, *t_im = NULL;
  int rc = PACG_OK;
  /* FP64 work buffers: the spectrum itself when stored in FP64 */
  if (sp == PACG_F64) {
    t_re = pl->spec_re;
    t_im = pl->spec_im;
  } else if (pacg_dev_malloc(&t_re, n * 8) || pacg_dev_malloc(&t_im, n * 8)) {
    rc = pacg_fail(PACG_ERR_NOMEM, "device allocation of %zu bytes failed", 16 * n);
    goto out;
  }
  const void *src_re = kr, *src_im = ki;
  if (!on_device) {
    if (pacg_memcpy_h2d(t_re, kr, n * 8, pl->stream) || pacg_memcpy_h2d(t_im, ki, n * 8, pl->stream)) {
      rc = pacg_fail(PACG_ERR_HIP, "host to device copy failed");
      goto out;
    }
    src_re = t_re;
    src_im = t_im;
  }
  if ((rc = run_all(pl, PACG_F64, PACG_K_DIF, src_re, src_im, t_re, t_im, 1.0))) goto out;
  if (sp != PACG_F64) {
    if (pacg_convert(t_re, t_im, PACG_F64, pl->spec_re, pl->spec_im, sp, n, 1.0, pl->stream)) {
      rc = pacg_fail(PACG_ERR_HIP, "spectrum conversion failed");
      goto out;
    }
  }
  if (pacg_stream_sync(pl->stream)) {
    rc = pacg_fail(PACG_ERR_HIP, "pacg_set_kernel: %s", pacg_hip_error_string(pacg_last_hip_error()));
    goto out;
  }
  pl->has_spec = 1;
out:
  if (sp != PACG_F64) {
    pacg_dev_free(t_re);
    pacg_dev_free(t_im);
  }
  return rc;
}

int pacg_get_spectrum(pacg_plan* pl, double* sr, double* si, int on_device) {
  if (!pl || !sr || !si) return pacg_fail(PACG_ERR_ARG, "NULL argument");
  if (!pl->has_spec) return pacg_fail(PACG_ERR_STATE, "no kernel spectrum set");
  size_t n = (size_t)pl->n;
  int sp = pl->d.spectrum_precision;
  if (sp == PACG_F64) {
    if (on_device) {
      HIPCHK(pacg_memcpy_d2d(sr, pl->spec_re, n * 8, pl->stream), "copy");
      HIPCHK(pacg_memcpy_d2d(si, pl->spec_im, n * 8, pl->stream), "copy");
    } else {
      HIPCHK(pacg_memcpy_d2h(sr, pl->spec_re, n * 8, pl->stream), "copy");
      HIPCHK(pacg_memcpy_d2h(si, pl->spec_im, n * 8, pl->stream), "copy");
    }
  } else {
    void *t_re = NULL, *t_im = NULL;
    if (on_device) {
      t_re = sr;
      t_im = si;
    } else if (pacg_dev_malloc(&t_re, n * 8) || pacg_dev_malloc(&t_im, n * 8)) {
      return pacg_fail(PACG_ERR_NOMEM, "device allocation failed");
    }
    HIPCHK(pacg_convert(pl->spec_re, pl->spec_im, sp, t_re, t_im, PACG_F64, n, 1.0, pl->stream), "convert");
    if (!on_device) {
      HIPCHK(pacg_memcpy_d2h(sr, t_re, n * 8, pl->stream), "copy");
      HIPCHK(pacg_memcpy_d2h(si, t_im, n * 8, pl->stream), "copy");
      pacg_stream_sync(pl->stream);
      pacg_dev_free(t_re);
      pacg_dev_free(t_im);
    }
  }
  HIPCHK(pacg_stream_sync(pl->stream), "stream synchronize");
  return PACG_OK;
}

int pacg_convolve(pacg_plan* pl, const void* xr, const void* xi, void* yr, void* yi) {
  if (!pl || !xr || !xi || !yr || !yi) return pacg_fail(PACG_ERR_ARG, "pacg_convolve: NULL argument");
  if (!pl->has_spec) return pacg_fail(PACG_ERR_STATE, "pacg_convolve: no kernel spectrum (pacg_set_kernel)");
  return run_conv(pl, pidx(pl), xr, xi, yr, yi, pl->spec_re, pl->spec_im, pl->d.spectrum_precision);
}

int pacg_convolve_pair(pacg_plan* pl, void* xr, void* xi, void* hr, void* hi) {
  if (!pl || !xr || !xi || !hr || !hi) return pacg_fail(PACG_ERR_ARG, "pacg_convolve_pair: NULL argument");
  int prec = pidx(pl);
  int rc = run_all(pl, prec, PACG_K_DIF, hr, hi, hr, hi, 1.0);
  if (rc) return rc;
  return run_conv(pl, prec, xr, xi, xr, xi, hr, hi, prec);
}

/* FFT-based convolution (convolution.c:12-29) with the scratch (sr, si) */
static int fft_based_with(pacg_plan* pl, void* xr, void* xi, void* hr, void* hi, void* sr, void* si) {
  int prec = pidx(pl), rc;
  if ((rc = execute_with(pl, PACG_OP_FFT, xr, xi, xr, xi, sr, si))) return rc;
  if ((rc = execute_with(pl, PACG_OP_FFT, hr, hi, hr, hi, sr, si))) return rc;
  HIPCHK(pacg_pointwise(xr, xi, hr, hi, prec, prec, (uint64_t)pl->n, 1.0, pl->stream), "pointwise");
  return execute_with(pl, PACG_OP_IFFT, xr, xi, xr, xi, sr, si);
}

int pacg_convolve_fft_based(pacg_plan* pl, void* xr, void* xi, void* hr, void* hi) {
  if (!pl || !xr || !xi || !hr || !hi) return pacg_fail(PACG_ERR_ARG, "pacg_convolve_fft_based: NULL argument");
  int rc;
  if ((rc = ensure_scratch(pl))) return rc;
  return fft_based_with(pl, xr, xi, hr, hi, pl->scr_re, pl->scr_im);
}

/* The same with the library's fixed device scratch: no allocation. */
int pacg_convolve_fft_based_noalloc(pacg_plan* pl, void* xr, void* xi, void* hr, void* hi) {
  if (!pl || !xr || !xi || !hr || !hi)
    return pacg_fail(PACG_ERR_ARG, "pacg_convolve_fft_based_noalloc: NULL argument");
  const size_t half = (size_t)pl->n * esize(pidx(pl));
  if (2 * half > (size_t)PACG_NOALLOC_SCRATCH_BYTES)
    return pacg_fail(PACG_ERR_SIZE, "pacg_convolve_fft_based_noalloc: %lld points need %zu bytes of scratch, more "
                     "than the fixed %u", (long long)pl->n, 2 * half, (unsigned)PACG_NOALLOC_SCRATCH_BYTES);
  void* base = NULL;
  HIPCHK(pacg_noalloc_scratch(&base), "scratch address");
  return fft_based_with(pl, xr, xi, hr, hi, base, (char*)base + half);
}

int pacg_graph_capture_convolve(pacg_plan* pl, const void* xr, const void* xi, void* yr, void* yi) {
  if (!pl) return pacg_fail(PACG_ERR_ARG, "NULL plan");
  if (!pl->has_spec) return pacg_fail(PACG_ERR_STATE, "no kernel spectrum set");
  if (pl->graph) {
    pacg_graph_destroy(pl->graph);
    pl->graph = NULL;
  }
  HIPCHK(pacg_graph_begin(pl->stream), "begin capture");
  int rc = pacg_convolve(pl, xr, xi, yr, yi);
  void* exec = NULL;
  int rc2 = pacg_graph_end(pl->stream, &exec);
  if (rc) return rc;
  HIPCHK(rc2, "end capture");
  pl->graph = exec;
  return PACG_OK;
}

int pacg_graph_replay(pacg_plan* pl) {
  if (!pl || !pl->graph) return pacg_fail(PACG_ERR_STATE, "no captured graph");
  HIPCHK(pacg_graph_launch(pl->graph, pl->stream), "graph launch");
  return PACG_OK;
}
