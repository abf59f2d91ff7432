// Batched permutation-avoiding convolutions of uneven power-of-two lengths
// (2^8 .. 2^14) in ONE persistent launch.
//
// Each pair (x_i, k_i) of length n_i gets the reference's convolution
// (convolution.c:31-72): s = A^T k, x <- A^T x, x <- (x .* s) / n, x <- conj(A) x,
// entirely inside one workgroup: k and x are read once, y is written once,
// the spectrum never leaves registers and no permutation is executed.
//
// Work is cut on the host into *items*: up to 2^14 / n convolutions of one
// length n (2^14 points per item), sorted longest first.  A workgroup of
// 1024 threads takes items blockIdx.x, blockIdx.x + gridDim.x, ... until none
// are left (every wave reaches the same exit).  Within an item, the n / E
// threads of one convolution each hold E points per register phase; the
// phases exchange through LDS (at most 2^14 points, 128 KB of (re, im) float
// pairs, one 8-byte access per point; slots swizzled by lds_swz).
// Schedules per length come from a variant table the autotuner picks from.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "pacg_batch.h"
#include "pacg_pass.hpp"

namespace pacg {

constexpr int BT = 1024;              // threads per workgroup
constexpr int BPTS = 1 << 14;         // points per item
constexpr int BLD = BPTS;             // LDS slots of (re, im) pairs (swizzled, no padding)

// Phase twiddle of phase PH.  The R points of a butterfly group take
// w^(mu * kr), kr = 0 .. R-1, with w = w_P^(2^sh); only the base b = w^mu
// comes from the two LDS tables of w_{2^14} (index e = (mu << sh) << (14 -
// log2 P); when the phase's stages span at most 128 points, e is a multiple
// of 128 and b = hi[e >> 7] exactly).  The other powers follow as products
// b^k = b^(k/2) * b^(k - k/2), at most three products deep: 6 complex
// multiplies instead of 7 table reads and products for a radix-8 group.
template <class SH, int PH, bool CJ>
__device__ __forceinline__ void batch_phase_twiddle(float* vr, float* vi, int tl, const float2* twh,
                                                    const float2* twl_) {
  constexpr int lo = SH::plo(PH);
  if constexpr (lo > 0) {
    constexpr int TW_SHIFT = 14 - SH::LGP;
    constexpr int R = 1 << SH::lgR(PH);
    constexpr int lgl = SH::lglam(lo);
    constexpr int sh = SH::LGP - SH::lglam(SH::phi(PH) + 1);
    constexpr bool ONE_TABLE = SH::lglam(SH::phi(PH) + 1) <= 7;
    static_assert(R <= 8, "phase radix");
    sfor<SH::E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      const int mu = (tl + SH::TPL * c) & ((1 << lgl) - 1);
      const int e = (mu << sh) << TW_SHIFT;
      float pr[R], pi[R];
      if constexpr (ONE_TABLE) {
        const float2 h = twh[e >> 7];
        pr[1] = h.x;
        pi[1] = h.y;
      } else {
        const float2 h = twh[e >> 7];
        const float2 l = twl_[e & 127];
        pr[1] = h.x * l.x - h.y * l.y;
        pi[1] = h.x * l.y + h.y * l.x;
      }
      if constexpr (CJ) pi[1] = -pi[1];
      sfor<R - 2>([&](auto k_) PACG_INL {
        constexpr int k = 2 + decltype(k_)::value;
        constexpr int a = k / 2, b = k - k / 2;
        pr[k] = pr[a] * pr[b] - pi[a] * pi[b];
        pi[k] = pr[a] * pi[b] + pi[a] * pr[b];
      });
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        constexpr int kr = SH::rev_phase(PH, r);
        if constexpr (kr != 0) cmul<float>(vr[c * R + r], vi[c * R + r], pr[kr], pi[kr]);
      });
    });
  }
}

template <class SH>
__device__ __forceinline__ void batch_item(const pacg_batch_args_t& a, const pacg_batch_item_t it, float2* lds,
                                           const float2* twh, const float2* twl_) {
  constexpr int P = SH::P, E = SH::E, TPL = SH::TPL, NP = SH::NP;
  static_assert(BT % TPL == 0, "threads per convolution (BT / TPL convolutions per item)");
  // Recompute every per-thread index inside the persistent loop: hoisted out
  // of it, the positions and LDS addresses of all phases would stay live.
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int g = tid / TPL;
  const int tl = tid % TPL;
  const bool valid = g < (int)it.count;
  const uint32_t base = valid ? a.conv_off[it.first + g] : 0u;
  // phase twiddles: w_P^idx = w_{2^14}^e from two LDS tables (batch_phase_twiddle)
  // LDS slot of (c, r) in phase PH: convolution g at g * P, swizzled (lds_swz)
  auto li = [&](auto PH_, int c, int r) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    return ((g * P) ^ lds_swz(pos<SH, PH>(tl, 0, 0))) ^ lds_swz(pos<SH, PH>(0, c, r));
  };

  // Spectrum A^T k and the read-ahead of x: in registers (at most 8 points per
  // thread below 2^13 points, so both fit beside the working points).
  static_assert(E <= 8 && P <= 4096, "register spectrum and x read-ahead");
  float vr[E], vi[E];  // working points
  float pr[E], pi[E];  // x, loaded while k is transformed
  float sr[E], si[E];  // the spectrum

  auto load_top = [&](const float* __restrict__ br, const float* __restrict__ bi, float* dr, float* di) PACG_INL {
    constexpr int PH = NP - 1;
    constexpr int R = 1 << SH::lgR(PH);
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int s = pos<SH, PH>(tl, c, r);
        dr[c * R + r] = valid ? br[base + s] : 0.f;
        di[c * R + r] = valid ? bi[base + s] : 0.f;
      });
    });
  };
  // Exchange number X of the item goes through LDS buffer X % 2 (an item holds
  // at most 1024 * 8 points, so two buffers fit): the buffer it writes was last
  // read two exchanges ago, before the barrier of the exchange in between, so
  // one barrier per exchange suffices.
  auto exchange = [&](auto FROM_, auto TO_, auto X_) PACG_INL {
    constexpr int FROM = decltype(FROM_)::value;
    constexpr int TO = decltype(TO_)::value;
    constexpr int RF = 1 << SH::lgR(FROM);
    constexpr int RT = 1 << SH::lgR(TO);
    constexpr int BUF = (decltype(X_)::value & 1) * (BPTS / 2);
    static_assert(BT * E <= BPTS / 2, "two exchange buffers per item");
    sfor<E / RF>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<RF>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int l = BUF + li(FROM_, c, r);
        lds[l] = make_float2(vr[c * RF + r], vi[c * RF + r]);
      });
    });
    __syncthreads();
    sfor<E / RT>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<RT>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int l = BUF + li(TO_, c, r);
        const float2 w = lds[l];
        vr[c * RT + r] = w.x;
        vi[c * RT + r] = w.y;
      });
    });
  };
  auto dif = [&](auto X0_) PACG_INL {  // X0: number of the first exchange
    sfor<NP>([&](auto i_) PACG_INL {
      constexpr int PH = NP - 1 - decltype(i_)::value;
      if constexpr (PH < NP - 1)
        exchange(std::integral_constant<int, PH + 1>{}, std::integral_constant<int, PH>{},
                 std::integral_constant<int, decltype(X0_)::value + NP - 2 - PH>{});
      phase_stages<float, SH, PH, M_DIF>(vr, vi);
      batch_phase_twiddle<SH, PH, false>(vr, vi, tl, twh, twl_);
    });
  };

  load_top(a.kr, a.ki, vr, vi);
  load_top(a.xr, a.xi, pr, pi);  // x arrives while k is transformed
  dif(std::integral_constant<int, 0>{});  // A^T k
#pragma unroll
  for (int e = 0; e < E; ++e) {
    sr[e] = vr[e];
    si[e] = vi[e];
    vr[e] = pr[e];
    vi[e] = pi[e];
  }
  dif(std::integral_constant<int, NP - 1>{});  // A^T x
  const float scale = 1.0f / (float)P;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    cmul<float>(vr[e], vi[e], sr[e], si[e]);
    vr[e] *= scale;
    vi[e] *= scale;
  }
  sfor<NP>([&](auto i_) PACG_INL {
    constexpr int PH = decltype(i_)::value;
    if constexpr (PH > 0)
      exchange(std::integral_constant<int, PH - 1>{}, std::integral_constant<int, PH>{},
               std::integral_constant<int, 2 * (NP - 1) + PH - 1>{});
    batch_phase_twiddle<SH, PH, true>(vr, vi, tl, twh, twl_);
    phase_stages<float, SH, PH, M_DITC>(vr, vi);
  });
  {
    constexpr int PH = NP - 1;
    constexpr int R = 1 << SH::lgR(PH);
    uint32_t sb = base;
    asm volatile("" : "+v"(sb));  // keep the store addresses from being formed early
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const uint32_t s = sb + (uint32_t)pos<SH, PH>(tl, c, r);
        if (valid) {
          a.yr[s] = vr[c * R + r];
          a.yi[s] = vi[c * R + r];
        }
      });
    });
  }
  __syncthreads();  // LDS is reused by the next item
}

// Lines of BT * 8 points and more (2^13, 2^14): an item holds G = 2^14 / P
// convolutions (2 at 2^13, 1 at 2^14) and each thread runs VT = G * TPL / BT
// virtual threads of 8 points in turn (virtual thread vtl: convolution
// vtl / TPL, thread vtl % TPL of it).  The lines stay in LDS between phases
// (the same LDS traffic as the exchange model, half the live registers).  The
// spectrum A^T k leaves the last forward phase of k in the registers of the
// thread that later holds the same positions of x (phase 0 of virtual thread
// v, for both), so it stays there: 2 x VT x 8 values, no global round trip.
template <class SH>
__device__ __forceinline__ void batch_item_vt(const pacg_batch_args_t& a, const pacg_batch_item_t it, float2* lds,
                                              const float2* twh, const float2* twl_) {
  constexpr int P = SH::P, E = SH::E, NP = SH::NP, TPL = SH::TPL;
  constexpr int G = BPTS / P;
  constexpr int VT = G * TPL / BT;
  static_assert(VT * BT == G * TPL && TPL % BT == 0 && E <= 8 && (G == 1 || G == 2), "virtual threads");
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  bool valid_g[G];
  uint32_t base_g[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    valid_g[g] = g < (int)it.count;
    base_g[g] = valid_g[g] ? a.conv_off[it.first + g] : 0u;
  }
  const float scale = 1.0f / (float)P;
  float sr[VT][E], si[VT][E];  // the spectrum at the phase-0 positions of each virtual thread

  // one phase for one virtual thread: read (global or LDS), compute, write back;
  // V = the virtual thread index when the phase touches the spectrum (tails 1, 2)
  auto run = [&](auto PH_, auto MODE_, auto FROMG_, auto TAIL_, auto V_, int vtl, const float* __restrict__ gr,
                 const float* __restrict__ gi) PACG_INL {
    constexpr int V = decltype(V_)::value;
    constexpr int PH = decltype(PH_)::value;
    constexpr int MODE = decltype(MODE_)::value;
    constexpr bool from_global = decltype(FROMG_)::value;
    constexpr int tail = decltype(TAIL_)::value;
    constexpr int R = 1 << SH::lgR(PH);
    // TPL >= BT: the convolution index is uniform over the workgroup
    const int g = G == 1 ? 0 : vtl / TPL;
    const int tl = G == 1 ? vtl : vtl % TPL;
    const bool valid = G == 1 ? valid_g[0] : (g ? valid_g[G - 1] : valid_g[0]);
    const uint32_t base = G == 1 ? base_g[0] : (g ? base_g[G - 1] : base_g[0]);
    const int lg = g * P;  // this convolution's LDS region
    float vr[E], vi[E];
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int sp = pos<SH, PH>(tl, c, r);
        if constexpr (from_global) {
          vr[c * R + r] = valid ? gr[base + sp] : 0.f;
          vi[c * R + r] = valid ? gi[base + sp] : 0.f;
        } else {
          const int l = lg + (lds_swz(pos<SH, PH>(tl, 0, 0)) ^ lds_swz(pos<SH, PH>(0, c, r)));
          const float2 w = lds[l];
          vr[c * R + r] = w.x;
          vi[c * R + r] = w.y;
        }
      });
    });
    if constexpr (MODE == M_DIF) {
      phase_stages<float, SH, PH, M_DIF>(vr, vi);
      batch_phase_twiddle<SH, PH, false>(vr, vi, tl, twh, twl_);
    } else {
      batch_phase_twiddle<SH, PH, true>(vr, vi, tl, twh, twl_);
      phase_stages<float, SH, PH, M_DITC>(vr, vi);
    }
    // tail: 0 write LDS, 1 k spectrum -> sr/si, 2 x: product + conj(A) phase 0 -> LDS, 3 -> y
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      if constexpr (tail == 2) {
        sfor<R>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          cmul<float>(vr[c * R + r], vi[c * R + r], sr[V][c * R + r], si[V][c * R + r]);
          vr[c * R + r] *= scale;
          vi[c * R + r] *= scale;
        });
      }
    });
    if constexpr (PH == 0) {
      if constexpr (tail == 2) phase_stages<float, SH, 0, M_DITC>(vr, vi);
    }
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int sp = pos<SH, PH>(tl, c, r);
        if constexpr (tail == 1) {
          sr[V][c * R + r] = vr[c * R + r];
          si[V][c * R + r] = vi[c * R + r];
        } else if constexpr (tail == 3) {
          if (valid) {
            a.yr[base + sp] = vr[c * R + r];
            a.yi[base + sp] = vi[c * R + r];
          }
        } else {
          const int l = lg + (lds_swz(pos<SH, PH>(tl, 0, 0)) ^ lds_swz(pos<SH, PH>(0, c, r)));
          lds[l] = make_float2(vr[c * R + r], vi[c * R + r]);
        }
      });
    });
  };

  // A^T k (the spectrum into sr/si), then A^T x, product, conj(A).  In the
  // forward phases the virtual threads run one after the other (one set of
  // points live; unrolled, they spill), except in phase 0, where each keeps its
  // own spectrum registers.  The conj(A) phases, with the spectrum gone, are
  // unrolled so that the LDS reads of one virtual thread overlap the other's
  // arithmetic.
  sfor<2>([&](auto pass_) PACG_INL {
    constexpr int pass = decltype(pass_)::value;
    const float* gr = pass == 0 ? a.kr : a.xr;
    const float* gi = pass == 0 ? a.ki : a.xi;
    sfor<NP>([&](auto i_) PACG_INL {
      constexpr int PH = NP - 1 - decltype(i_)::value;
      constexpr int tail = PH > 0 ? 0 : (pass == 0 ? 1 : 2);
      __syncthreads();
      if constexpr (tail == 0) {
#pragma nounroll
        for (int v = 0; v < VT; ++v)
          run(std::integral_constant<int, PH>{}, std::integral_constant<int, M_DIF>{},
              std::integral_constant<bool, PH == NP - 1>{}, std::integral_constant<int, tail>{},
              std::integral_constant<int, 0>{}, tid + v * BT, gr, gi);
      } else {
        sfor<VT>([&](auto v_) PACG_INL {
          constexpr int v = decltype(v_)::value;
          run(std::integral_constant<int, PH>{}, std::integral_constant<int, M_DIF>{},
              std::integral_constant<bool, PH == NP - 1>{}, std::integral_constant<int, tail>{},
              std::integral_constant<int, v>{}, tid + v * BT, gr, gi);
        });
      }
    });
  });
  sfor<NP - 1>([&](auto i_) PACG_INL {
    constexpr int PH = 1 + decltype(i_)::value;
    constexpr int tail = PH == NP - 1 ? 3 : 0;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < VT; ++v)
      run(std::integral_constant<int, PH>{}, std::integral_constant<int, M_DITC>{}, std::integral_constant<bool, false>{},
          std::integral_constant<int, tail>{}, std::integral_constant<int, 0>{}, tid + v * BT, a.xr, a.xi);
  });
  __syncthreads();  // LDS is reused by the next item
}

// Variant table: (log2 n, variant) -> shape.  Stage 0 first; phases listed
// by their top stage.  Kept in step with pacg_batch_variants[] below.
#define PACG_BATCH_VARIANTS(X)                                          \
  X(8, 0, 3, 0x2Fu, 3, 0x210u)        /* 8 | 8 | 4             */     \
  X(8, 1, 4, 0xAAu, 4, 0x3210u)       /* 4 | 4 | 4 | 4         */     \
  X(9, 0, 3, 0x3Fu, 3, 0x210u)        /* 8 | 8 | 8             */     \
  X(9, 1, 4, 0xABu, 4, 0x3210u)       /* 8 | 4 | 4 | 4         */     \
  X(10, 0, 4, 0x7Fu, 4, 0x3210u)      /* 8 | 8 | 8 | 2         */     \
  X(10, 1, 4, 0xAFu, 4, 0x3210u)      /* 8 | 8 | 4 | 4         */     \
  X(11, 0, 4, 0xBFu, 4, 0x3210u)      /* 8 | 8 | 8 | 4         */     \
  X(11, 1, 5, 0x2F9u, 4, 0x4321u)     /* 2 4 | 8 | 8 | 4       */     \
  X(12, 0, 4, 0xFFu, 4, 0x3210u)      /* 8 | 8 | 8 | 8         */     \
  X(13, 0, 5, 0x1FFu, 5, 0x43210u)    /* 8 | 8 | 8 | 8 | 2     */     \
  X(14, 0, 5, 0x3FEu, 5, 0x43210u)    /* 4 | 8 | 8 | 8 | 8 (two 8-point virtual threads) */

__global__ void __launch_bounds__(BT) batch_kernel(const pacg_batch_args_t a) {
  __shared__ float2 lds[BLD];
  __shared__ float2 twh[128], twl[128];
  {
    const float2* tw = reinterpret_cast<const float2*>(a.tw);
    const int t = threadIdx.x;
    if (t < 128) twh[t] = tw[t * 128];
    else if (t < 256) twl[t - 128] = tw[t - 128];
  }
  __syncthreads();
  for (uint32_t i = blockIdx.x; i < a.n_items; i += gridDim.x) {
    const pacg_batch_item_t it = a.items[i];
    const int key = (int)it.log2n * 4 + (int)it.variant;
    switch (key) {
#define PACG_CASE(M, V, NS, SC, NPH, PC)                                    \
  case M * 4 + V:                                                          \
    static_assert(Shape<NS, SC, NPH, PC>::LGP == M, "variant length");    \
    if constexpr (Shape<NS, SC, NPH, PC>::P >= BT * 8)                     \
      batch_item_vt<Shape<NS, SC, NPH, PC>>(a, it, lds, twh, twl);    \
    else                                                                   \
      batch_item<Shape<NS, SC, NPH, PC>>(a, it, lds, twh, twl);       \
    break;
      PACG_BATCH_VARIANTS(PACG_CASE)
#undef PACG_CASE
      default:
        break;
    }
  }
}

}  // namespace pacg

extern "C" {

// Stage schedules of the variants (stage 0 first), for the host and tests.
#define PACG_SH(NS, SC, NPH, PC) pacg::Shape<NS, SC, NPH, PC>
#define PACG_RHO(NS, SC, NPH, PC, G) (NS > G ? PACG_SH(NS, SC, NPH, PC)::rho(G) : 0)
#define PACG_ROW(M, V, NS, SC, NPH, PC)                                                                     \
  {M, V, NS,                                                                                             \
   {PACG_RHO(NS, SC, NPH, PC, 0), PACG_RHO(NS, SC, NPH, PC, 1), PACG_RHO(NS, SC, NPH, PC, 2),            \
    PACG_RHO(NS, SC, NPH, PC, 3), PACG_RHO(NS, SC, NPH, PC, 4), PACG_RHO(NS, SC, NPH, PC, 5),            \
    PACG_RHO(NS, SC, NPH, PC, 6)},                                                                       \
   PACG_SH(NS, SC, NPH, PC)::TPL},
const pacg_batch_variant_t pacg_batch_variants[] = {PACG_BATCH_VARIANTS(PACG_ROW)};
#undef PACG_RHO
#undef PACG_SH
#undef PACG_ROW
const int pacg_num_batch_variants = sizeof(pacg_batch_variants) / sizeof(pacg_batch_variants[0]);

int pacg_batch_launch(const pacg_batch_args_t* a, int blocks, void* stream) {
  if (a->n_items == 0) return 0;
  hipLaunchKernelGGL(pacg::batch_kernel, dim3(blocks), dim3(pacg::BT), 0, (hipStream_t)stream, *a);
  return (int)hipGetLastError();
}

int pacg_cu_count(int* cus) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  return (int)hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
}

}  // extern "C"
