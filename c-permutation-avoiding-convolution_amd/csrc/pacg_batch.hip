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
// phases exchange through LDS (at most 2^14 points, 135 KB with padding).
// Schedules per length come from a variant table the autotuner picks from.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "pacg_batch.h"
#include "pacg_pass.hpp"

namespace pacg {

constexpr int BT = 1024;              // threads per workgroup
constexpr int BPTS = 1 << 14;         // points per item
constexpr int BLD = BPTS + BPTS / 32; // padded LDS length per array

template <class SH>
__device__ __forceinline__ void batch_item(const pacg_batch_args_t& a, const pacg_batch_item_t it, float* lre,
                                           float* lim) {
  constexpr int P = SH::P, E = SH::E, TPL = SH::TPL, NP = SH::NP;
  static_assert(BT % TPL == 0, "threads per convolution");
  constexpr int G = BT / TPL;  // convolution slots per item
  constexpr int LDP = P + P / 32;
  const int tid = threadIdx.x;
  const int g = tid / TPL;
  const int tl = tid % TPL;
  const bool valid = g < (int)it.count;
  const uint32_t base = valid ? a.conv_off[it.first + g] : 0u;
  const float2* __restrict__ tw = reinterpret_cast<const float2*>(a.tw);
  constexpr int TW_SHIFT = 14 - SH::LGP;  // w_P^idx = w_{2^14}^{idx << TW_SHIFT}
  auto twl = [&](int idx, float& wr, float& wi) PACG_INL {
    const float2 w = tw[idx << TW_SHIFT];
    wr = w.x;
    wi = w.y;
  };
  auto li = [&](int s) PACG_INL { return g * LDP + s + (s >> 5); };

  float vr[E], vi[E];   // working points
  float pr[E], pi[E];   // x prefetched while k is transformed
  float sr[E], si[E];   // spectrum A^T k (phase-0 layout)

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
  auto exchange = [&](auto FROM_, auto TO_) PACG_INL {
    constexpr int FROM = decltype(FROM_)::value;
    constexpr int TO = decltype(TO_)::value;
    constexpr int RF = 1 << SH::lgR(FROM);
    constexpr int RT = 1 << SH::lgR(TO);
    __syncthreads();
    sfor<E / RF>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<RF>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int l = li(pos<SH, FROM>(tl, c, r));
        lre[l] = vr[c * RF + r];
        lim[l] = vi[c * RF + r];
      });
    });
    __syncthreads();
    sfor<E / RT>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<RT>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int l = li(pos<SH, TO>(tl, c, r));
        vr[c * RT + r] = lre[l];
        vi[c * RT + r] = lim[l];
      });
    });
  };
  auto dif = [&]() PACG_INL {
    sfor<NP>([&](auto i_) PACG_INL {
      constexpr int PH = NP - 1 - decltype(i_)::value;
      if constexpr (PH < NP - 1) exchange(std::integral_constant<int, PH + 1>{}, std::integral_constant<int, PH>{});
      phase_stages<float, SH, PH, M_DIF>(vr, vi);
      phase_twiddle_f<float, SH, PH, false>(vr, vi, tl, twl);
    });
  };

  load_top(a.kr, a.ki, vr, vi);
  load_top(a.xr, a.xi, pr, pi);
  dif();  // A^T k
#pragma unroll
  for (int e = 0; e < E; ++e) {
    sr[e] = vr[e];
    si[e] = vi[e];
    vr[e] = pr[e];
    vi[e] = pi[e];
  }
  dif();  // A^T x
  const float scale = 1.0f / (float)P;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    cmul<float>(vr[e], vi[e], sr[e], si[e]);
    vr[e] *= scale;
    vi[e] *= scale;
  }
  sfor<NP>([&](auto i_) PACG_INL {
    constexpr int PH = decltype(i_)::value;
    if constexpr (PH > 0) exchange(std::integral_constant<int, PH - 1>{}, std::integral_constant<int, PH>{});
    phase_twiddle_f<float, SH, PH, true>(vr, vi, tl, twl);
    phase_stages<float, SH, PH, M_DITC>(vr, vi);
  });
  {
    constexpr int PH = NP - 1;
    constexpr int R = 1 << SH::lgR(PH);
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int s = pos<SH, PH>(tl, c, r);
        if (valid) {
          a.yr[base + s] = vr[c * R + r];
          a.yi[base + s] = vi[c * R + r];
        }
      });
    });
  }
  __syncthreads();  // LDS is reused by the next item
}

// Variant table: (log2 n, variant) -> shape.  Stage 0 first; phases listed
// by their top stage.  Kept in step with pacg_batch_variants[] below.
#define PACG_BATCH_VARIANTS(X)                                    \
  X(8, 0, 4, 0xAAu, 2, 0x31u)         /* 4 4 | 4 4      */ \
  X(8, 1, 3, 0x2Fu, 3, 0x210u)        /* 8 | 8 | 4      */ \
  X(9, 0, 3, 0x3Fu, 3, 0x210u)        /* 8 | 8 | 8      */ \
  X(9, 1, 5, 0x2A9u, 3, 0x431u)       /* 2 4 | 4 4 | 4  */ \
  X(10, 0, 5, 0x2AAu, 3, 0x431u)      /* 4 4 | 4 4 | 4  */ \
  X(10, 1, 4, 0xFDu, 3, 0x321u)       /* 2 8 | 8 | 8    */ \
  X(10, 2, 4, 0xAFu, 3, 0x310u)       /* 8 | 8 | 4 4    */ \
  X(11, 0, 6, 0x6AAu, 3, 0x531u)      /* 4 4 | 4 4 | 4 2 */ \
  X(11, 1, 4, 0xBFu, 4, 0x3210u)      /* 8 | 8 | 8 | 4  */ \
  X(12, 0, 6, 0xAAAu, 3, 0x531u)      /* 4 4 | 4 4 | 4 4 */ \
  X(12, 1, 4, 0xFFu, 4, 0x3210u)      /* 8 | 8 | 8 | 8  */ \
  X(13, 0, 7, 0x1AAAu, 4, 0x6531u)    /* 4 4 |4 4 |4 4 | 2 */ \
  X(13, 1, 5, 0x1FFu, 4, 0x4210u)     /* 8 | 8 | 8 | 8 2 */ \
  X(14, 0, 7, 0x2AAAu, 4, 0x6531u)    /* 4 4 |4 4 |4 4 | 4 */ \
  X(14, 1, 5, 0x3FDu, 4, 0x4321u)     /* 2 8 | 8 | 8 | 8 */

__global__ void __launch_bounds__(BT) batch_kernel(const pacg_batch_args_t a) {
  __shared__ float lre[BLD];
  __shared__ float lim[BLD];
  for (uint32_t i = blockIdx.x; i < a.n_items; i += gridDim.x) {
    const pacg_batch_item_t it = a.items[i];
    const int key = (int)it.log2n * 4 + (int)it.variant;
    switch (key) {
#define PACG_CASE(M, V, NS, SC, NPH, PC)                \
  case M * 4 + V:                                      \
    batch_item<Shape<NS, SC, NPH, PC>>(a, it, lre, lim); \
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
