/* Batched uneven-length convolutions: shared between pacg_batch.hip (the
 * persistent kernel) and pacg_batch.c (items, autotuner, graphs).  Plain C. */
#ifndef PACG_BATCH_H
#define PACG_BATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* up to 1024 / TPL convolutions of one length n (TPL threads each), at sorted
 * positions [first, first + count) */
typedef struct {
  uint32_t first;
  uint16_t count;
  uint8_t log2n;
  uint8_t variant;
} pacg_batch_item_t;

typedef struct {
  const float* xr;
  const float* xi;
  const float* kr;
  const float* ki;
  float* yr;
  float* yi;
  const uint32_t* conv_off;          /* element offset of the pair at sorted position i */
  const pacg_batch_item_t* items;    /* grouped by workgroup, each group ordered by shape */
  const uint32_t* wg_first;          /* workgroup w runs items [wg_first[w], wg_first[w + 1]) */
  uint32_t n_items;
  const void* tw;                    /* w_{2^14}^p, p < 2^14, (re, im) float pairs */
  uint32_t total;                    /* points in each array (< 2^30) */
} pacg_batch_args_t;

typedef struct {
  int log2n, variant, nstages;
  int radix[7];                      /* stage 0 first */
  int tpl;                           /* threads per convolution */
  int cost;                          /* time of one item on one workgroup, 0.1 us (MI355X) */
  int cost_mixed;                    /* the same in a batch of mixed lengths (fitted), for the slice dealing */
} pacg_batch_variant_t;

extern const pacg_batch_variant_t pacg_batch_variants[];
extern const int pacg_num_batch_variants;

int pacg_batch_launch(const pacg_batch_args_t* a, int blocks, void* stream);
int pacg_cu_count(int* cus);

#ifdef __cplusplus
}
#endif

#endif
