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
// radix-2/4 stages of the same run on the VALU, the pass-external twiddle /
// spectral product on the VALU, LDS -> global.
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
  // at most 2^127: mx < 2^128 keeps the scaled values below 2 (inside FP16), and 2^128 would not
  // be a finite FP32 factor to scale back with
  ex = ex > 127 ? 127 : ex;
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

// Stage list of an MFMA pass: NS stages, 2 bits each (log2 radix, stage 0 first)
template <int NS, uint32_t SC>
struct MShape {
  static constexpr int lg(int g) { return (int)((SC >> (2 * g)) & 3u); }
  static constexpr int rad(int g) { return 1 << lg(g); }
  static constexpr int lglam(int g) {
    int v = 0;
    for (int h = 0; h < g; ++h) v += lg(h);
    return v;
  }
  static constexpr int LGP = lglam(NS);
  static constexpr int P = 1 << LGP;
  // mixed digit reversal of a point index over the pass's stages
  static __device__ __forceinline__ int rev(int s) {
    int out = 0;
#pragma unroll
    for (int g = 0; g < NS; ++g) out |= ((s >> lglam(g)) & (rad(g) - 1)) << (LGP - lglam(g + 1));
    return out;
  }
};

// KIND: DIF (A^T), DIT (A), DITC (conj A), FUSED.  Radix-8 stages run on the
// matrix cores, radix-2/4 stages of the same pass on the VALU, all on one LDS
// tile of LINES lines.
template <typename TS, int NS, uint32_t SC, int KIND, int LAYOUT, int LINES>
__global__ void __launch_bounds__(mf::NT) mfma_pass_kernel(const pacg_pass_args_t a) {
  using MS = MShape<NS, SC>;
  constexpr int P = MS::P;
  constexpr int LDL = P + 1;  // LDS row per line (+1 against bank conflicts)
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

  // stage G over the tile: DIF (F then twiddle), DIT / DITC (twiddle then F);
  // leg m of butterfly (mu, beta) at beta * R lam + mu + m lam, twiddle
  // w_{R lam}^{mu m} = w_P table entry mu m P / (R lam)
  auto stage = [&](auto G_, int mode) {
    constexpr int G = decltype(G_)::value;
    constexpr int R = MS::rad(G);
    constexpr int lam = 1 << MS::lglam(G);
    constexpr int NBF = LINES * (P / R);  // butterflies of the stage in the tile
    if constexpr (R == 8) {
      static_assert(NBF % 16 == 0, "whole MFMA groups");
      constexpr int NGRP = NBF / 16;
      for (int grp = wave; grp < NGRP; grp += mf::NT / 64) {
        const int bf = grp * 16 + jl;  // butterfly id in the tile
        const int ll = bf / (P / 8), b = bf % (P / 8);
        const int mu = b % lam, beta = b / lam;
        const int s0 = beta * (8 * lam) + mu;
        float* arr = is_im ? tim : tre;
        float x[4];
        for (int v = 0; v < 4; ++v) x[v] = arr[ll * LDL + s0 + (leg0 + v) * lam];
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
    } else {
      for (int bf = tid; bf < NBF; bf += mf::NT) {
        const int ll = bf / (P / R), b = bf % (P / R);
        const int mu = b % lam, beta = b / lam;
        const int s0 = ll * LDL + beta * (R * lam) + mu;
        float xr[R], xi[R];
#pragma unroll
        for (int m = 0; m < R; ++m) {
          xr[m] = tre[s0 + m * lam];
          xi[m] = tim[s0 + m * lam];
        }
        auto twiddle = [&](bool cj) {
#pragma unroll
          for (int m = 1; m < R; ++m) {
            const int idx = (mu * m) * (P / (R * lam));
            cmul<float>(xr[m], xi[m], twr[idx], cj ? -twi[idx] : twi[idx]);
          }
        };
        if (mode == M_DIF) {
          dft<float, R, false, 0, 1>(xr, xi);
          if (lam > 1) twiddle(false);
        } else {
          if (lam > 1) twiddle(mode == M_DITC);
          if (mode == M_DITC)
            dft<float, R, true, 0, 1>(xr, xi);
          else
            dft<float, R, false, 0, 1>(xr, xi);
        }
#pragma unroll
        for (int m = 0; m < R; ++m) {
          tre[s0 + m * lam] = xr[m];
          tim[s0 + m * lam] = xi[m];
        }
      }
    }
  };

  // element-wise step over the tile: pass-external twiddle, spectral product
  auto elementwise = [&](int what) {
    for (int idx = tid; idx < LINES * P; idx += mf::NT) {
      const int ll = idx / P, s = idx % P;
      bool ok;
      uint32_t j;
      const uint32_t base = line_base(ll, ok, j);
      float vr = tre[ll * LDL + s], vi = tim[ll * LDL + s];
      if (what == 0 || what == 1) {  // external twiddle w_K^{j' k(s)} (conj when what == 1)
        const uint32_t e = ((j >> a.log2_inner) * (uint32_t)MS::rev(s)) << a.ext_shift;
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
    sfor<NS>([&](auto i_) {
      stage(std::integral_constant<int, NS - 1 - decltype(i_)::value>{}, M_DIF);
      __syncthreads();
    });
    if (KIND == PACG_K_DIF && a.ext) {
      elementwise(0);
      __syncthreads();
    }
    if (KIND == PACG_K_FUSED) {
      elementwise(2);
      __syncthreads();
      sfor<NS>([&](auto i_) {
        stage(i_, M_DITC);
        __syncthreads();
      });
    }
  } else {
    const int mode = KIND == PACG_K_DITC ? M_DITC : M_DIT;
    if (a.ext) {
      elementwise(mode == M_DITC ? 1 : 0);
      __syncthreads();
    }
    sfor<NS>([&](auto i_) {
      stage(i_, mode);
      __syncthreads();
    });
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

template <typename TS, int NS, uint32_t SC, int KIND, int LAYOUT, int LINES>
int launch_mfma(const pacg_pass_args_t* a, void* stream) {
  const uint64_t lines = LAYOUT == PACG_L_STRIDED ? ((uint64_t)a->n_o << a->log2_S) : (uint64_t)a->n_o;
  const uint32_t blocks = (uint32_t)((lines + LINES - 1) / LINES);
  if (blocks == 0) return 0;
  hipLaunchKernelGGL((mfma_pass_kernel<TS, NS, SC, KIND, LAYOUT, LINES>), dim3(blocks), dim3(mf::NT), 0,
                     (hipStream_t)stream, *a);
  return (int)hipGetLastError();
}

}  // namespace pacg

using namespace pacg;

// one registry row: (stages, stage code, lines per tile, rank, P, kind, layout, spectrum type, name)
#define MF_ROW(NS, SC, LINES, RANK, P, KIND, LAYOUT, TS, SP, NAME)                                            \
  {PACG_F32, SP, KIND, LAYOUT, NS, SC, 0u, P, LINES, mf::NT, RANK, &launch_mfma<TS, NS, SC, KIND, LAYOUT, LINES>, \
   NAME, 1}
// every kind and layout of one shape and tile (rank 0: the default tile, > 0: autotuner alternatives)
#define MF_SHAPE(NS, SC, LINES, RANK, P, NM)                                                                   \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_DIF, PACG_L_STRIDED, float, PACG_F32, "mfma_f32_dif_s_" NM),           \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_DIT, PACG_L_STRIDED, float, PACG_F32, "mfma_f32_dit_s_" NM),           \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_DITC, PACG_L_STRIDED, float, PACG_F32, "mfma_f32_ditc_s_" NM),         \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_FUSED, PACG_L_STRIDED, float, PACG_F32, "mfma_f32_fused_s_" NM "_h"),  \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_FUSED, PACG_L_STRIDED, double, PACG_F64, "mfma_f32_fused_s_" NM "_h64"), \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_DIF, PACG_L_CONTIG, float, PACG_F32, "mfma_f32_dif_c_" NM),            \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_DIT, PACG_L_CONTIG, float, PACG_F32, "mfma_f32_dit_c_" NM),            \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_DITC, PACG_L_CONTIG, float, PACG_F32, "mfma_f32_ditc_c_" NM),          \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_FUSED, PACG_L_CONTIG, float, PACG_F32, "mfma_f32_fused_c_" NM "_h"),   \
  MF_ROW(NS, SC, LINES, RANK, P, PACG_K_FUSED, PACG_L_CONTIG, double, PACG_F64, "mfma_f32_fused_c_" NM "_h64")

extern "C" {
extern pacg_variant_t pacg_mfma_variants[];
// runs that hold at least one radix-8 stage, up to 512 points (stage codes: 2 bits per
// stage, log2 radix, stage 0 lowest); tiles of at most 4096 points keep four or more
// workgroups resident per CU
pacg_variant_t pacg_mfma_variants[] = {
    MF_SHAPE(1, 0x3u, 128, 0, 8, "8_l128"),
    MF_SHAPE(1, 0x3u, 64, 1, 8, "8_l64"),
    MF_SHAPE(2, 0xFu, 32, 0, 64, "8x8_l32"),
    MF_SHAPE(2, 0xFu, 16, 1, 64, "8x8_l16"),
    MF_SHAPE(3, 0x3Fu, 8, 0, 512, "8x8x8_l8"),
    MF_SHAPE(3, 0x3Fu, 4, 1, 512, "8x8x8_l4"),
    MF_SHAPE(3, 0x2Fu, 8, 0, 256, "8x8x4_l8"),
    MF_SHAPE(3, 0x2Fu, 16, 1, 256, "8x8x4_l16"),
    MF_SHAPE(3, 0x3Eu, 8, 0, 256, "4x8x8_l8"),
    MF_SHAPE(3, 0x3Eu, 16, 1, 256, "4x8x8_l16"),
    MF_SHAPE(2, 0xBu, 32, 0, 32, "8x4_l32"),
    MF_SHAPE(2, 0xEu, 32, 0, 32, "4x8_l32"),
    MF_SHAPE(2, 0x7u, 64, 0, 16, "8x2_l64"),
    MF_SHAPE(2, 0xDu, 64, 0, 16, "2x8_l64"),
};
extern const int pacg_num_mfma_variants;
const int pacg_num_mfma_variants = sizeof(pacg_mfma_variants) / sizeof(pacg_mfma_variants[0]);
}
