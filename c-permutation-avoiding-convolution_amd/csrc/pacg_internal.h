/*
 * Internal C ABI between the C host layer (pacg_plan.c, pacg_dropin.c) and
 * the HIP kernel objects (pacg_pass_*.hip, pacg_aux.hip).  Plain C: no HIP
 * or torch types cross this line, only pointers, sizes and a stream handle
 * passed as void*.
 */
#ifndef PACG_INTERNAL_H
#define PACG_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision of a working array */
enum { PACG_F64 = 0, PACG_F32 = 1 };

/* what a pass applies to its lines (a "line" = the P points of one
 * sub-problem, spaced S elements apart) */
enum {
  PACG_K_DIF = 0,   /* A^T stages, top to bottom  (transposed butterflies)  */
  PACG_K_DIT = 1,   /* A stages, bottom to top    (butterflies)             */
  PACG_K_DITC = 2,  /* conj(A) stages             (conjugate butterflies)   */
  PACG_K_FUSED = 3  /* A^T stages, * spectrum * scale, conj(A) stages       */
};

enum { PACG_L_STRIDED = 0, PACG_L_CONTIG = 1 };

/* Arguments of one pass launch.  Element addresses: line (o, j) point s is
 * o * P * S + s * S + j, with j in [0, S) and o in [0, n_o).  For the
 * contiguous layout S = 1 and n_o counts lines. */
typedef struct {
  const void* in_re;
  const void* in_im;
  void* out_re;
  void* out_im;
  const void* spec_re;   /* fused passes: spectrum at the same positions */
  const void* spec_im;
  const void* tw_p;      /* w_P^p, p < P, interleaved (re, im), working precision */
  const void* tw_hi;     /* axis table, two levels: w_N^e = hi[e >> lo_bits] * lo[e & mask] */
  const void* tw_lo;
  uint32_t log2_S;
  uint32_t n_o;
  uint32_t log2_inner;   /* j' (within-axis lower index) = j >> log2_inner */
  uint32_t ext_shift;    /* log2(N_axis / K), K = P * L */
  uint32_t lo_bits;
  uint32_t ext;          /* 1: apply the pass-external twiddle w_K^{j' k} */
  uint32_t vec16;        /* 1: in/out base pointers are 16-byte aligned (vector access allowed) */
  uint32_t spec_nt;      /* fused, strided, FP64 spectrum: load it non-temporally (array + spectrum exceed the Infinity Cache) */
  uint32_t in_nt;        /* A^T pass reading the caller's x (first launch of a convolution): load it non-temporally */
  double scale;          /* fused: factor of the spectral product (1/n), applied at the store */
  double out_scale;      /* factor applied at the final store (1 or 1/n) */
} pacg_pass_args_t;

typedef int (*pacg_launch_fn)(const pacg_pass_args_t* a, void* stream);

/* One compiled pass kernel. */
typedef struct {
  int prec;              /* PACG_F64 / PACG_F32 (data) */
  int spec_prec;         /* spectrum precision for fused passes */
  int kind;              /* PACG_K_* */
  int layout;            /* PACG_L_* */
  int nstages;
  uint32_t stage_code;   /* 2 bits per stage: log2(radix), stage 0 lowest */
  uint32_t phase_code;   /* 4 bits per phase: index of its top stage */
  int P;
  int lines;             /* lines per workgroup (TJ or NL) */
  int threads;
  int rank;              /* 0: the default tile for its key; > 0: autotuner alternatives */
  pacg_launch_fn launch;
  const char* name;
  int engine;            /* 0: VALU butterflies, 1: radix-8 stages on MFMA (split FP16) */
} pacg_variant_t;

/* registry of every compiled pass kernel (generated: gen/pacg_variants_index.c) */
extern const pacg_variant_t* const pacg_variant_tables[];
extern const int pacg_variant_counts[];
extern const int pacg_num_variant_tables;
/* radix-8 MFMA pass kernels (pacg_mfma.hip) */
extern pacg_variant_t pacg_mfma_variants[];
extern const int pacg_num_mfma_variants;

/* ---- auxiliary kernels (pacg_aux.hip) ---------------------------------- */
/* K1: twiddle tables generated in FP64 on the device.
 *   tw_p:  w_P^p, p < P          (P entries)
 *   hi/lo: two-level table of w_N (N/2^lo_bits and 2^lo_bits entries)
 * Written as interleaved pairs in the requested precision (rounded once). */
int pacg_gen_twiddles(void* out, int prec, uint32_t N, uint32_t count, uint32_t step, void* stream);

/* K7: digit-reversal gather / scatter over an nd array, rho from the axis
 * schedules (same closed form as the host helper). out[p] = in[rho(p)] (gather)
 * or out[rho(p)] = in[p] (scatter).  Not in place. */
typedef struct {
  int nd;
  uint32_t dims[3];
  int nst[3];
  uint8_t radix_log2[3][32];
} pacg_perm_desc_t;
int pacg_permute(const void* in_re, const void* in_im, void* out_re, void* out_im, int prec,
                 const pacg_perm_desc_t* d, int scatter, void* stream);

/* table gather out[p] = in[rho[p]] with a caller-supplied rho (device int array) */
int pacg_gather_rho(const void* in_re, const void* in_im, void* out_re, void* out_im, const int* rho, uint64_t n,
                    int prec, void* stream);
/* pointwise z <- (z * w) * scale over n complex (FFT-based convolution) */
int pacg_pointwise(void* zre, void* zim, const void* wre, const void* wim, int prec, int wprec,
                   uint64_t n, double scale, void* stream);
/* y <- x * scale, converting precision (also a plain copy when scale == 1) */
int pacg_convert(const void* xre, const void* xim, int xprec, void* yre, void* yim, int yprec,
                 uint64_t n, double scale, void* stream);

/* Device address, on the current device, of the fixed scratch array of the
 * allocation-free entry point (PACG_NOALLOC_SCRATCH_BYTES, include/pacg.h);
 * resolved with hipGetSymbolAddress once per device.  Returns a HIP error code. */
int pacg_noalloc_scratch(void** p);

/* HBM probe: mode 0 read `bytes` of src, 1 write `bytes` to dst, 2 copy src -> dst
 * (16-byte accesses; the read mode writes dst[0] only in an impossible case). */
int pacg_bw_probe(int mode, void* dst, const void* src, uint64_t bytes, void* stream);

/* runtime glue implemented in HIP (so the C layer needs no HIP headers) */
int pacg_dev_malloc(void** p, size_t bytes);
int pacg_dev_free(void* p);
int pacg_host_malloc(void** p, size_t bytes); /* pinned host memory (hipHostMalloc) */
int pacg_host_free(void* p);
int pacg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int pacg_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int pacg_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int pacg_stream_sync(void* stream);
int pacg_stream_create(void** stream);
int pacg_stream_destroy(void* stream);
int pacg_device_sync(void);
int pacg_get_device(int* dev);
int pacg_device_count(int* n);
int pacg_num_cus(void); /* compute units of the current device (cached per device) */
int pacg_last_hip_error(void);
const char* pacg_hip_error_string(int code);
int pacg_event_create(void** ev);
int pacg_event_destroy(void* ev);
int pacg_event_record(void* ev, void* stream);
int pacg_event_sync(void* ev);
int pacg_event_elapsed_ms(float* ms, void* start, void* stop);
int pacg_graph_begin(void* stream);
int pacg_graph_end(void* stream, void** exec);
int pacg_graph_launch(void* exec, void* stream);
int pacg_graph_destroy(void* exec);

#ifdef __cplusplus
}
#endif

#endif /* PACG_INTERNAL_H */
