/*
 * pacg.h -- MI355X (gfx950) permutation-avoiding FFT convolution, extended C API.
 *
 * The drop-in entry points with the reference's own names live in
 * pacg_fft.h; this header is the plan-based interface beneath them: device
 * buffers, explicit mixed-radix schedules, FP64 or FP32 working precision,
 * a cached kernel spectrum, error codes and a HIP stream.  No HIP or torch
 * types appear in any signature; streams are passed as void*.
 *
 * Layout: split complex (separate re / im arrays), row-major, dims[nd-1]
 * contiguous -- the reference's layout (butterflies_2d.c:23, butterflies_3d.c:25).
 *
 * Operators (reference notation, fft.h:17-231):
 *   A        DIT butterflies            apply_butterflies_r*          (butterflies.c:12)
 *   conj(A)  conjugate butterflies      apply_conjugate_butterflies_r* (butterflies.c:41)
 *   A^T      transposed butterflies     apply_transposed_butterflies_r*(butterflies.c:70)
 * with A P = F (DFT) for the schedule's digit reversal P, so A^T x = P F x is
 * the DFT in digit-reversed order.  The convolution
 *   y = (1/n) conj(A) [ (A^T k) .* (A^T x) ]          (convolution.c:31-72)
 * never executes P.
 */
#ifndef PACG_H
#define PACG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------- */
enum {
  PACG_OK = 0,
  PACG_ERR_ARG = 1,       /* invalid argument (dims, radix, schedule, NULL) */
  PACG_ERR_SIZE = 2,      /* total size >= 2^31 or an axis not equal to its schedule product */
  PACG_ERR_NOMEM = 3,     /* host or device allocation failed */
  PACG_ERR_HIP = 4,       /* a HIP runtime call failed (message names it) */
  PACG_ERR_NOKERNEL = 5,  /* no compiled pass kernel covers the requested schedule */
  PACG_ERR_STATE = 6      /* e.g. convolve before a kernel spectrum was set */
};

enum { PACG_PREC_F64 = 0, PACG_PREC_F32 = 1 };

/* operations of pacg_execute (reference fft.c:12-88 and butterflies.c) */
enum {
  PACG_OP_TRANS = 0,  /* A^T            (transposed butterflies; fft2 without P) */
  PACG_OP_CONJ = 1,   /* conj(A)        (conjugate butterflies)                  */
  PACG_OP_FWD = 2,    /* A              (butterflies; ufft)                      */
  PACG_OP_FFT = 3,    /* A P            (fft: natural-order DFT)                 */
  PACG_OP_FFT2 = 4,   /* P^-1 A^T       (fft2: natural-order DFT)                */
  PACG_OP_IFFT = 5,   /* conj(A) P / n  (ifft)                                   */
  PACG_OP_UIFFT = 6,  /* conj(A) / n    (uifft)                                  */
  PACG_OP_PERM = 7    /* P: x[p] <- x[rho(p)], the digit-reversal permutation of
                         apply_permutation_split (index_reversals.c:85-94); in
                         place through the plan's scratch                         */
};

#define PACG_MAX_STAGES 32

/* Plan description.  Axis a has length dims[a] = prod radix[a][0..nstages[a]-1];
 * nstages[a] = 0 asks for the default schedule (radix-4 stages with one
 * radix-2 stage on top when log2 dims[a] is odd).  Stage 0 has unit stride;
 * its radix is the innermost digit of the digit reversal.  nstages[0] = -1
 * makes axis 0 a batch axis: any length, not transformed (independent
 * convolutions of the rows, each scaled by 1 / its own size). */
typedef struct {
  int nd;                                 /* 1, 2 or 3 */
  int dims[3];                            /* each a power of two, total < 2^31 */
  int nstages[3];
  int radix[3][PACG_MAX_STAGES];          /* entries in {2, 4, 8} */
  int precision;                          /* PACG_PREC_F64 or PACG_PREC_F32: data arrays */
  int spectrum_precision;                 /* storage of A^T k (always computed in FP64) */
  int fused_axis;                         /* axis whose last pass carries the product; -1 = the
                                             first axis with stages, or the last one when the
                                             first one's would run strided lines of more than
                                             512 points */
  int mfma;                               /* FP32 plans: run radix-8 stages as F_8 (x) I contractions
                                             on the matrix cores (split FP16, FP32 accumulation) */
} pacg_desc;

typedef struct pacg_plan pacg_plan;

/* Fill a descriptor with pure radix-r schedules (r in 2, 4, 8). */
int pacg_desc_init(pacg_desc* d, int nd, const int* dims, int r, int precision);

int pacg_plan_create(pacg_plan** plan, const pacg_desc* d);
int pacg_plan_destroy(pacg_plan* plan);
/* Stream for every later call (default: the plan's own non-blocking stream). */
int pacg_plan_set_stream(pacg_plan* plan, void* stream);
void* pacg_plan_stream(const pacg_plan* plan);
long long pacg_plan_size(const pacg_plan* plan);
/* Digit-reversal map of the plan's schedules: rho[p] = natural index held at p. */
int pacg_plan_rho(const pacg_plan* plan, int* rho);
/* Number of kernel launches of one pacg_convolve and a text summary of them. */
int pacg_plan_describe(const pacg_plan* plan, char* buf, int buflen);

/* Device-buffer transforms in the plan precision.  in == out is allowed. */
int pacg_execute(pacg_plan* plan, int op, const void* in_re, const void* in_im, void* out_re, void* out_im);

/* Kernel spectrum: computes A^T k in FP64 on the device from FP64 input (host
 * or device memory) and stores it in spectrum_precision. */
int pacg_set_kernel(pacg_plan* plan, const double* k_re, const double* k_im, int k_on_device);
/* Copy the stored spectrum (unscaled A^T k) to FP64 device or host memory. */
int pacg_get_spectrum(pacg_plan* plan, double* s_re, double* s_im, int on_device);

/* y = (1/n) conj(A)[(A^T k) .* (A^T x)] with the stored spectrum, device buffers. */
int pacg_convolve(pacg_plan* plan, const void* x_re, const void* x_im, void* y_re, void* y_im);
/* Reference semantics on device buffers: h <- A^T h (unscaled), x <- convolution. */
int pacg_convolve_pair(pacg_plan* plan, void* x_re, void* x_im, void* h_re, void* h_im);
/* FFT-based (permuting) convolution: h <- F h, x <- ifft(F x .* F h) (convolution.c:12-29). */
int pacg_convolve_fft_based(pacg_plan* plan, void* x_re, void* x_im, void* h_re, void* h_im);

/* Allocation-free variant of pacg_convolve_fft_based, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * kernels and results as pacg_convolve_fft_based; the digit-reversal scratch
 * is a fixed device array built into the library instead of plan memory
 * allocated on first use: PACG_NOALLOC_SCRATCH_BYTES on every GPU the library
 * runs on, held until the process ends.  A plan whose scratch (2 x elements x
 * element size: 16 bytes per FP64 point, 8 per FP32) exceeds it is rejected
 * with PACG_ERR_SIZE.  All calls on one GPU share that array, and it keeps the
 * last call's contents: calls that overlap on one GPU (several streams or
 * threads at once) race.  The call itself makes no hipMalloc / hipFree. */
#define PACG_NOALLOC_SCRATCH_BYTES (32u << 20)
int pacg_convolve_fft_based_noalloc(pacg_plan* plan, void* x_re, void* x_im, void* h_re, void* h_im);

/* Capture pacg_convolve(x -> y) on the plan stream into a HIP graph and replay it. */
int pacg_graph_capture_convolve(pacg_plan* plan, const void* x_re, const void* x_im, void* y_re, void* y_im);
int pacg_graph_replay(pacg_plan* plan);

/* Autotuner: times every compiled tile shape of each launch of pacg_convolve
 * on these device buffers (x -> y, `reps` launches each) and keeps the
 * fastest per launch; then times whole convolutions with the first launch (an
 * A^T pass reading x) loading x by default or non-temporally, and keeps the
 * faster policy (non-temporal: x, read once, leaves the Infinity Cache to y
 * and the spectrum; pacg_plan_describe then ends in " x:nt").  Returns how many
 * launches changed kernel or load policy, or minus an error code.  Drops a
 * captured graph (capture again afterwards). */
int pacg_plan_autotune(pacg_plan* plan, const void* x_re, const void* x_im, void* y_re, void* y_im, int reps);

/* Radix-schedule search: for the shape, precisions and switches of d, tries
 * the candidate schedules of each transformed axis in turn with the other axes
 * fixed (two sweeps), each as a plan with its tiles autotuned, timing `reps`
 * convolutions of seeded random data on scratch buffers it allocates, and
 * writes the fastest schedules into d->nstages / d->radix (a schedule replaces
 * the current one only when it is more than 1 % faster).  When d->fused_axis
 * is -1 it then tries each transformed axis for the fused pass and writes the
 * fastest into d->fused_axis (-1, the default rule, unless an axis is more
 * than 1 % faster).  Candidates of a
 * length 2^t: every mix of radix-8, -4 and at most one radix-2 stage, each with
 * its radices descending and ascending from stage 0 (pacg_schedule_candidates).
 * The caller's schedules (defaults where nstages is 0) are the starting point
 * and must make a plan.  Returns how many axes changed schedule, or -error. */
int pacg_plan_search(pacg_desc* d, int reps);
/* The candidate schedules of a length n = 2^t, one per row of `out`
 * (max_stages ints, stage 0 first, the unused tail 0): writes up to max_cand
 * rows and returns how many candidates there are, or -error. */
int pacg_schedule_candidates(int n, int* out, int max_cand, int max_stages);

/* Tile shapes of launch i (0-based, in launch order) of pacg_convolve: writes
 * up to max_names kernel names (static strings) of the compiled kernels that
 * can run it and returns how many there are, or minus an error code.
 * pacg_plan_pin_launch makes the named
 * one current, as the autotuner would; it drops a captured graph. */
int pacg_plan_launch_candidates(pacg_plan* plan, int launch, const char** names, int max_names);
int pacg_plan_pin_launch(pacg_plan* plan, int launch, const char* name);
/* Sets the load policy of the first launch of pacg_convolve, as the autotuner
 * would: non-temporal loads of x when `on` is nonzero (effective where that
 * launch is an A^T pass kernel on the VALU, that is where a convolution has two
 * or more launches and its first pass has no MFMA stage).  Drops a captured
 * graph. */
int pacg_plan_stream_input(pacg_plan* plan, int on);

/* Timing aid: runs `reps` convolutions with a HIP event before every kernel
 * launch (on the plan stream) and writes the mean duration of launch i to
 * launch_ms[i] and its kernel name to names[i] (if names != NULL).  Returns
 * the number of launches per convolution, or minus an error code. */
int pacg_profile_convolve(pacg_plan* plan, const void* x_re, const void* x_im, void* y_re, void* y_im, int reps,
                          float* launch_ms, const char** names, int max_launches);

/* ---- batched convolutions of uneven lengths (one persistent launch) --------
 * count pairs (x_i, k_i) of lengths 2^m_i, 8 <= m_i <= 14, fewer than 2^30 points in
 * all, FP32 split complex,
 * packed back to back in input order (pair i at pacg_batch_offset(b, i) in each
 * of the six arrays).  y_i = (1/n_i) conj(A_i)[(A_i^T k_i) .* (A_i^T x_i)] with a
 * per-length schedule chosen from compiled variants (pacg_batch_autotune). */
typedef struct pacg_batch pacg_batch;
int pacg_batch_create(pacg_batch** batch, int count, const int* log2_len);
int pacg_batch_destroy(pacg_batch* batch);
long long pacg_batch_total(const pacg_batch* batch);
long long pacg_batch_offset(const pacg_batch* batch, int i);
void* pacg_batch_stream(const pacg_batch* batch);
int pacg_batch_items(const pacg_batch* batch);
int pacg_batch_convolve(pacg_batch* batch, const float* x_re, const float* x_im, const float* k_re, const float* k_im,
                        float* y_re, float* y_im);
/* Chooses a compiled schedule per length present and the dealing of items to
 * workgroups (below): for each dealing, each length's schedules are timed on
 * the whole batch (these buffers), the other lengths at their current
 * schedules, over two sweeps, keeping the fastest; then the faster dealing
 * with its schedules is kept.  Returns how many lengths changed schedule, or
 * -error (the batch then keeps its schedules and dealing).  Drops a captured
 * graph. */
int pacg_batch_autotune(pacg_batch* batch, const float* x_re, const float* x_im, const float* k_re,
                        const float* k_im, float* y_re, float* y_im, int reps);
int pacg_batch_set_variant(pacg_batch* batch, int log2n, int variant);
int pacg_batch_get_variant(const pacg_batch* batch, int log2n);
/* Dealing of the work items to the workgroups of the launch: longest
 * processing time first (the default), or contiguous slices of the
 * shape-ordered item list (one or two shapes per workgroup: fewer run starts,
 * balanced by fitted costs).  Results do not depend on it.  Set drops a
 * captured graph; get returns the dealing or -error. */
enum { PACG_DEAL_LPT = 0, PACG_DEAL_SLICES = 1 };
int pacg_batch_set_dealing(pacg_batch* batch, int dealing);
int pacg_batch_get_dealing(const pacg_batch* batch);
int pacg_batch_num_variants(int log2n);
/* Stage radices (stage 0 first) of a variant; returns the stage count or -error. */
int pacg_batch_schedule(int log2n, int variant, int* radix, int max_stages);
int pacg_batch_capture(pacg_batch* batch, const float* x_re, const float* x_im, const float* k_re, const float* k_im,
                       float* y_re, float* y_im);
int pacg_batch_replay(pacg_batch* batch);
int pacg_batch_sync(pacg_batch* batch);

/* Free the device plans and buffers cached by the drop-in fft.h entry points. */
void pacg_dropin_release(void);

/* HBM probe, the device counterpart of the reference's host bandwidth tool
 * (memory_bw_test.c:27-67, read / write / copy of large arrays): mode 0 reads
 * `bytes` of src, 1 writes `bytes` to dst, 2 copies src to dst, with 16-byte
 * accesses from a grid that fills the device.  `reps` launches are timed with
 * HIP events on `stream` (NULL: the default stream); *ms receives the mean time
 * of one launch.  bytes must be a positive multiple of 16 and src / dst device
 * buffers of at least that size (src unused by mode 1, dst by mode 0). */
int pacg_hbm_probe(int mode, void* dst, const void* src, uint64_t bytes, int reps, void* stream, float* ms);

/* Launch floor: mean time per launch, in microseconds, of an empty kernel
 * launched `reps` times back to back on `stream` (NULL: a private stream) and
 * of a HIP graph holding one such launch replayed `reps` times.  Measurement
 * aid; no counterpart in the reference. */
int pacg_launch_floor(int reps, void* stream, float* direct_us, float* graph_us);

/* Wait for the plan stream. */
int pacg_sync(pacg_plan* plan);

/* Last error message of the calling thread ("" when none). */
const char* pacg_last_error(void);
/* Library version string. */
const char* pacg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PACG_H */
