// Radix-8 stages on the matrix cores: F_8 (x) I as batched dense 8x8 complex
// contractions, split-FP16 operands, FP32 accumulation (gfx950 MFMA).
//
// A radix-8 stage (reference butterflies.c:277-389 A, :391-503 conj A,
// :505-632 A^T) applies F_8 to every butterfly's 8 legs.  Written in real
// arithmetic one butterfly is y = M x with the 16 x 16 matrix
//   M = [[Re F, -Im F], [Im F, Re F]],  x = (Re x_0..7, Im x_0..7).
// Sixteen butterflies form the B operand of v_mfma_f32_16x16x16f16: lane
// (q = lane / 16, j = lane % 16) holds the four values 4q..4q+3 of the
// stacked vector of butterfly j, i.e. Re legs 0-3, Re 4-7, Im 0-3, Im 4-7 for
// q = 0..3; the result D comes back in the same arrangement.
//
// Split FP16: x = hi + lo 2^-11 with hi = fp16(x), lo = fp16((x - hi) 2^11),
// likewise M; D = M_hi x_hi + 2^-11 (M_hi x_lo + M_lo x_hi) with FP32
// accumulation (the lo*lo term is below FP32 rounding).  Each butterfly is
// scaled by a power of two (its largest magnitude) before the split and back
// after, so FP16's range never limits the data.
//
// The pass (one read and one write of HBM, like pacg_pass.hpp) stages its
// tile of lines through LDS: global -> LDS, one MFMA phase per radix-8 stage
// (twiddles on the VALU, re/im partners exchanged between lanes l and l^32),
// the pass-external twiddle / spectral product on the VALU, LDS -> global.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "pacg_internal.h"
#include "pacg_pass.hpp"

namespace pacg {
namespace mf {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int NT = 256;  // threads per workgroup (4 waves)

// entry (i, k) of M for F_8 (INV: conjugate), exp(-2 pi i a b / 8)
__device__ __forceinline__ float mentry(int i, int k, bool inv) {
  const int a = i & 7, b = k & 7;
  const int e = (a * b) & 7;
  const float h = 0.70710678118654752440f;
  const float C[8] = {1.f, h, 0.f, -h, -1.f, -h, 0.f, h};
  const float S[8] = {0.f, -h, -1.f, -h, 0.f, h, 1.f, h};
  const float re = C[e];
  const float im = inv ? -S[e] : S[e];
  const bool ro = i < 8, co = k < 8;
  if (ro && co) return re;
  if (ro && !co) return -im;
  if (!ro && co) return im;
  return re;
}

struct Mat {
  half4_t hi, lo;
};

__device__ __forceinline__ Mat make_mat(int lane, bool inv) {
  Mat m;
  const int i = lane & 15;
  for (int v = 0; v < 4; ++v) {
    const int k = 4 * (lane >> 4) + v;
    const float x = mentry(i, k, inv);
    const _Float16 h = (_Float16)x;
    m.hi[v] = h;
    m.lo[v] = (_Float16)((x - (float)h) * 2048.0f);
  }
  return m;
}

// y = M x for the 16 butterflies of one group; x, y: this lane's 4 values
__device__ __forceinline__ void f8_mfma(const Mat& m, float x[4]) {
  // power-of-two scale from the butterfly's largest magnitude (lanes j, j+16, j+32, j+48)
  float mx = fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3])));
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  int ex = 0;
  if (mx > 0.f) ex = ((int)((__float_as_uint(mx) >> 23) & 0xff)) - 126;  // mx = f 2^ex, f in [0.5, 1)
  const float down = ldexpf(1.0f, -ex), up = ldexpf(1.0f, ex);
  half4_t bh, bl;
  for (int v = 0; v < 4; ++v) {
    const float s = x[v] * down;
    const _Float16 h = (_Float16)s;
    bh[v] = h;
    bl[v] = (_Float16)((s - (float)h) * 2048.0f);
  }
  float4_t z = {0.f, 0.f, 0.f, 0.f};
  float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x16f16(m.hi, bh, z, 0, 0, 0);
  float4_t d2 = __builtin_amdgcn_mfma_f32_16x16x16f16(m.hi, bl, z, 0, 0, 0);
  d2 = __builtin_amdgcn_mfma_f32_16x16x16f16(m.lo, bh, d2, 0, 0, 0);
  for (int v = 0; v < 4; ++v) x[v] = (d1[v] + d2[v] * (1.0f / 2048.0f)) * up;
}

}  // namespace mf

// KIND: DIF (A^T), DIT (A), DITC (conj A), FUSED; NS8 radix-8 stages (P = 8^NS8)
template <typename TS, int NS8, int KIND, int LAYOUT, int LINES>
__global__ void __launch_bounds__(mf::NT) mfma_pass_kernel(const pacg_pass_args_t a) {
  constexpr int P = NS8 == 1 ? 8 : 64;
  constexpr int LDL = P + 1;                 // LDS row per line (+1 against bank conflicts)
  constexpr int NBF = LINES * (P / 8);       // butterflies per stage in the tile
  static_assert(NBF % 16 == 0, "whole MFMA groups");
  constexpr int NGRP = NBF / 16;
  __shared__ float tre[LINES * LDL], tim[LINES * LDL];
  __shared__ float twr[P], twi[P];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t S = 1u << a.log2_S;
  const uint64_t nlines = (LAYOUT == PACG_L_STRIDED) ? ((uint64_t)a.n_o << a.log2_S) : (uint64_t)a.n_o;
  {
    const float2* tp = reinterpret_cast<const float2*>(a.tw_p);
    for (int i = tid; i < P; i += mf::NT) {
      const float2 w = tp[i];
      twr[i] = w.x;
      twi[i] = w.y;
    }
  }
  // line ll of the tile -> element offset of its point 0, and its j (for the external twiddle)
  auto line_base = [&](int ll, bool& ok, uint32_t& j) -> uint32_t {
    const uint64_t lg = (uint64_t)tile * LINES + ll;
    ok = lg < nlines;
    if (LAYOUT == PACG_L_STRIDED) {
      const uint32_t o = (uint32_t)(lg >> a.log2_S);
      j = (uint32_t)lg & (S - 1u);
      return o * (uint32_t)P * S + j;
    }
    j = 0;
    return (uint32_t)lg * (uint32_t)P;
  };
  const uint32_t stride = (LAYOUT == PACG_L_STRIDED) ? S : 1u;
  const float* __restrict__ inr = reinterpret_cast<const float*>(a.in_re);
  const float* __restrict__ ini = reinterpret_cast<const float*>(a.in_im);
  float* __restrict__ outr = reinterpret_cast<float*>(a.out_re);
  float* __restrict__ outi = reinterpret_cast<float*>(a.out_im);

  // global -> LDS (strided: consecutive threads walk consecutive lines)
  for (int idx = tid; idx < LINES * P; idx += mf::NT) {
    int ll, s;
    if (LAYOUT == PACG_L_STRIDED) {
      ll = idx % LINES;
      s = idx / LINES;
    } else {
      ll = idx / P;
      s = idx % P;
    }
    bool ok;
    uint32_t j;
    const uint32_t off = line_base(ll, ok, j) + (uint32_t)s * stride;
    tre[ll * LDL + s] = ok ? inr[off] : 0.f;
    tim[ll * LDL + s] = ok ? ini[off] : 0.f;
  }
  __syncthreads();

  const mf::Mat mfwd = mf::make_mat(lane, false);
  const mf::Mat minv = mf::make_mat(lane, true);
  const int q = lane >> 4, jl = lane & 15;
  const bool is_im = q >= 2;
  const int leg0 = 4 * (q & 1);

  // one radix-8 stage g over the tile: DIF (F then twiddle), DIT / DITC (twiddle then F)
  auto stage = [&](int g, int mode) {
    const int lam = g == 0 ? 1 : 8;  // sub-stride of stage g within the line
    for (int grp = wave; grp < NGRP; grp += mf::NT / 64) {
      const int bf = grp * 16 + jl;            // butterfly id in the tile
      const int ll = bf / (P / 8), b = bf % (P / 8);
      const int mu = b % lam, beta = b / lam;
      const int s0 = beta * (8 * lam) + mu;    // leg m at s0 + m * lam
      float* arr = is_im ? tim : tre;
      float x[4];
      for (int v = 0; v < 4; ++v) x[v] = arr[ll * LDL + s0 + (leg0 + v) * lam];
      // twiddle w_{8 lam}^{mu m} (index in the w_P table: mu m P / (8 lam))
      auto twiddle = [&](bool cj) {
        for (int v = 0; v < 4; ++v) {
          const float other = __shfl_xor(x[v], 32);  // partner lane: the other component
          const int idx = (mu * (leg0 + v)) * (P / (8 * lam));
          const float c = twr[idx];
          const float sn = cj ? -twi[idx] : twi[idx];
          const float re = is_im ? other : x[v], im = is_im ? x[v] : other;
          x[v] = is_im ? (re * sn + im * c) : (re * c - im * sn);
        }
      };
      if (mode == M_DIF) {
        mf::f8_mfma(mfwd, x);
        if (lam > 1) twiddle(false);
      } else {
        if (lam > 1) twiddle(mode == M_DITC);
        mf::f8_mfma(mode == M_DITC ? minv : mfwd, x);
      }
      for (int v = 0; v < 4; ++v) arr[ll * LDL + s0 + (leg0 + v) * lam] = x[v];
    }
  };

  // element-wise step over the tile: pass-external twiddle, spectral product, final scale
  auto elementwise = [&](int what) {
    for (int idx = tid; idx < LINES * P; idx += mf::NT) {
      const int ll = idx / P, s = idx % P;
      bool ok;
      uint32_t j;
      const uint32_t base = line_base(ll, ok, j);
      float vr = tre[ll * LDL + s], vi = tim[ll * LDL + s];
      if (what == 0 || what == 1) {  // external twiddle w_K^{j' k(s)} (conj when what == 1)
        int ks = 0;  // digit reversal of s over the pass's radix-8 stages
        ks = NS8 == 1 ? s : ((s & 7) << 3) | (s >> 3);
        const uint32_t e = ((j >> a.log2_inner) * (uint32_t)ks) << a.ext_shift;
        const float2 h = reinterpret_cast<const float2*>(a.tw_hi)[e >> a.lo_bits];
        const float2 l = reinterpret_cast<const float2*>(a.tw_lo)[e & ((1u << a.lo_bits) - 1u)];
        float wr = h.x * l.x - h.y * l.y, wi = h.x * l.y + h.y * l.x;
        if (what == 1) wi = -wi;
        cmul<float>(vr, vi, wr, wi);
      } else if (what == 2) {  // spectral product and 1/n
        const TS* __restrict__ hr = reinterpret_cast<const TS*>(a.spec_re);
        const TS* __restrict__ hi = reinterpret_cast<const TS*>(a.spec_im);
        const uint32_t off = base + (uint32_t)s * stride;
        const float wr = ok ? (float)hr[off] : 0.f, wi = ok ? (float)hi[off] : 0.f;
        cmul<float>(vr, vi, wr, wi);
        vr *= (float)a.scale;
        vi *= (float)a.scale;
      }
      tre[ll * LDL + s] = vr;
      tim[ll * LDL + s] = vi;
    }
  };

  if (KIND == PACG_K_DIF || KIND == PACG_K_FUSED) {
    for (int g = NS8 - 1; g >= 0; --g) {
      stage(g, M_DIF);
      __syncthreads();
    }
    if (KIND == PACG_K_DIF && a.ext) {
      elementwise(0);
      __syncthreads();
    }
    if (KIND == PACG_K_FUSED) {
      elementwise(2);
      __syncthreads();
      for (int g = 0; g < NS8; ++g) {
        stage(g, M_DITC);
        __syncthreads();
      }
    }
  } else {
    const int mode = KIND == PACG_K_DITC ? M_DITC : M_DIT;
    if (a.ext) {
      elementwise(mode == M_DITC ? 1 : 0);
      __syncthreads();
    }
    for (int g = 0; g < NS8; ++g) {
      stage(g, mode);
      __syncthreads();
    }
  }

  // LDS -> global
  const float osc = (float)a.out_scale;
  for (int idx = tid; idx < LINES * P; idx += mf::NT) {
    int ll, s;
    if (LAYOUT == PACG_L_STRIDED) {
      ll = idx % LINES;
      s = idx / LINES;
    } else {
      ll = idx / P;
      s = idx % P;
    }
    bool ok;
    uint32_t j;
    const uint32_t off = line_base(ll, ok, j) + (uint32_t)s * stride;
    if (ok) {
      outr[off] = tre[ll * LDL + s] * osc;
      outi[off] = tim[ll * LDL + s] * osc;
    }
  }
}

template <typename TS, int NS8, int KIND, int LAYOUT, int LINES>
int launch_mfma(const pacg_pass_args_t* a, void* stream) {
  const uint64_t lines = LAYOUT == PACG_L_STRIDED ? ((uint64_t)a->n_o << a->log2_S) : (uint64_t)a->n_o;
  const uint32_t blocks = (uint32_t)((lines + LINES - 1) / LINES);
  if (blocks == 0) return 0;
  hipLaunchKernelGGL((mfma_pass_kernel<TS, NS8, KIND, LAYOUT, LINES>), dim3(blocks), dim3(mf::NT), 0,
                     (hipStream_t)stream, *a);
  return (int)hipGetLastError();
}

}  // namespace pacg

using namespace pacg;

#define MF_ROW(NS8, KIND, LAYOUT, LINES, TS, SP, NAME)                                                        \
  {PACG_F32, SP, KIND, LAYOUT, NS8, NS8 == 1 ? 0x3u : 0xFu, NS8 == 1 ? 0x0u : 0x10u, NS8 == 1 ? 8 : 64, LINES, \
   mf::NT, 0, &launch_mfma<TS, NS8, KIND, LAYOUT, LINES>, NAME, 1}

extern "C" {
extern pacg_variant_t pacg_mfma_variants[];
pacg_variant_t pacg_mfma_variants[] = {
    MF_ROW(2, PACG_K_DIF, PACG_L_STRIDED, 32, float, PACG_F32, "mfma_f32_dif_s_8x8"),
    MF_ROW(2, PACG_K_DIT, PACG_L_STRIDED, 32, float, PACG_F32, "mfma_f32_dit_s_8x8"),
    MF_ROW(2, PACG_K_DITC, PACG_L_STRIDED, 32, float, PACG_F32, "mfma_f32_ditc_s_8x8"),
    MF_ROW(2, PACG_K_FUSED, PACG_L_STRIDED, 32, float, PACG_F32, "mfma_f32_fused_s_8x8_h"),
    MF_ROW(2, PACG_K_FUSED, PACG_L_STRIDED, 32, double, PACG_F64, "mfma_f32_fused_s_8x8_h64"),
    MF_ROW(2, PACG_K_DIF, PACG_L_CONTIG, 32, float, PACG_F32, "mfma_f32_dif_c_8x8"),
    MF_ROW(2, PACG_K_DIT, PACG_L_CONTIG, 32, float, PACG_F32, "mfma_f32_dit_c_8x8"),
    MF_ROW(2, PACG_K_DITC, PACG_L_CONTIG, 32, float, PACG_F32, "mfma_f32_ditc_c_8x8"),
    MF_ROW(2, PACG_K_FUSED, PACG_L_CONTIG, 32, float, PACG_F32, "mfma_f32_fused_c_8x8_h"),
    MF_ROW(2, PACG_K_FUSED, PACG_L_CONTIG, 32, double, PACG_F64, "mfma_f32_fused_c_8x8_h64"),
    MF_ROW(1, PACG_K_DIF, PACG_L_STRIDED, 128, float, PACG_F32, "mfma_f32_dif_s_8"),
    MF_ROW(1, PACG_K_DIT, PACG_L_STRIDED, 128, float, PACG_F32, "mfma_f32_dit_s_8"),
    MF_ROW(1, PACG_K_DITC, PACG_L_STRIDED, 128, float, PACG_F32, "mfma_f32_ditc_s_8"),
    MF_ROW(1, PACG_K_FUSED, PACG_L_STRIDED, 128, float, PACG_F32, "mfma_f32_fused_s_8_h"),
    MF_ROW(1, PACG_K_FUSED, PACG_L_STRIDED, 128, double, PACG_F64, "mfma_f32_fused_s_8_h64"),
    MF_ROW(1, PACG_K_DIF, PACG_L_CONTIG, 128, float, PACG_F32, "mfma_f32_dif_c_8"),
    MF_ROW(1, PACG_K_DIT, PACG_L_CONTIG, 128, float, PACG_F32, "mfma_f32_dit_c_8"),
    MF_ROW(1, PACG_K_DITC, PACG_L_CONTIG, 128, float, PACG_F32, "mfma_f32_ditc_c_8"),
    MF_ROW(1, PACG_K_FUSED, PACG_L_CONTIG, 128, float, PACG_F32, "mfma_f32_fused_c_8_h"),
    MF_ROW(1, PACG_K_FUSED, PACG_L_CONTIG, 128, double, PACG_F64, "mfma_f32_fused_c_8_h64"),
};
extern const int pacg_num_mfma_variants;
const int pacg_num_mfma_variants = sizeof(pacg_mfma_variants) / sizeof(pacg_mfma_variants[0]);
}
