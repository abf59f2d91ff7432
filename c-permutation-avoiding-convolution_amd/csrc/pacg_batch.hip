// Batched permutation-avoiding convolutions of uneven power-of-two lengths
// (2^8 .. 2^14) in ONE persistent launch.
//
// Each pair (x_i, k_i) of length n_i gets the reference's convolution
// (convolution.c:31-72): s = A^T k, x <- A^T x, x <- (x .* s) / n, x <- conj(A) x,
// entirely inside one workgroup: k and x are read once, y is written once,
// the spectrum never leaves registers and no permutation is executed.
//
// Work is cut on the host into *items*: 1024 / TPL convolutions of one length
// n, where the TPL = n / E threads of one convolution hold E (8 or 16) points
// each per register phase.  The host deals the items to the workgroups
// (longest processing time first) and orders each workgroup's list by shape,
// so a workgroup meets a run of items of one shape: while it transforms the
// k of one item it has the x of that item in flight, and during its conj(A)
// phases the k of the next item of the run.  The loads are unconditional
// (a missing convolution of a ragged item reads the item's first pair and
// stores nothing), so they stay in flight until their first use.  Global
// accesses go through buffer resources: a 32-bit per-thread byte offset plus
// a compile-time offset per element (scalar), bounded by the array size.
// Phases exchange points through LDS: (re, im) float pairs, one 8-byte access
// per point, one pad slot per 16 points (per 8 for 8-point shapes), so a slot
// is a per-thread base plus a compile-time offset; an item of 16384 points
// uses one buffer (two barriers per exchange), one of 8192 points alternates
// two.
#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include "pacg_batch.h"
#include "pacg_pass.hpp"

namespace pacg {

static_assert(sizeof(pacg_batch_item_t) == 8 && offsetof(pacg_batch_item_t, count) == 4 &&
                  offsetof(pacg_batch_item_t, log2n) == 6 && offsetof(pacg_batch_item_t, variant) == 7,
              "item record read as two dwords");

constexpr int BT = 1024;              // threads per workgroup
constexpr int BPTS = 1 << 14;         // points of the largest item (BT x 16)
constexpr int BLD = BPTS + BPTS / 8;  // LDS slots of (re, im) pairs: two 8192-point buffers padded 1 in 8

// LDS slot of point s of an exchange tile padded with one slot per 2^PS
// points.  For a sum of two positions with disjoint bits (a per-thread part
// and a compile-time part) the slot is the sum of their slots, so the
// compile-time part lands in the instruction's offset field.  16 consecutive
// lanes of an 8-byte access then meet 16 different slots modulo 16 in every
// phase layout of the 16-point shapes (pad 1 in 16: runs of consecutive
// points, or points 16 apart for the bottom phase, whose slots step by 17)
// and of the 8-point shapes (pad 1 in 8: points 8 or 64 apart step by 9 or 72).
template <int PS>
__device__ __forceinline__ constexpr int lds_pad(int s) {
  return s + (s >> PS);
}

using rsrc_t = __amdgpu_buffer_rsrc_t;
// raw buffer resource over `bytes` bytes (out-of-range loads read 0, stores
// drop); the last argument of the accesses is the cache policy, 2 = nt
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_load(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void buf_store(float v, rsrc_t r, uint32_t voff, uint32_t soff) {
  // non-temporal: y is not read again by this launch (about 6 % faster than
  // the default policy on the 4096-pair batch)
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 2);
}

// Group index o of group c of thread tl in phase PH.  The top phase (the one
// that meets global memory) deals groups out as tl + TPL * c, so the lanes of
// a wave touch consecutive points; the other phases keep a thread's groups
// adjacent, tl * (E / R) + c, so a thread's points stay in a block of E * tl
// points and the exchanges below the top phase can stay inside a wavefront.
template <class SH, int PH>
__device__ __forceinline__ int group_of(int tl, int c) {
  constexpr int lgg = SH::lgE() - SH::lgR(PH);  // log2 groups per thread
  if constexpr (PH == SH::NP - 1)
    return tl + SH::TPL * c;
  else
    return (tl << lgg) | c;
}
// Position (within the line) of element (c, r) of thread tl in phase PH: the
// phase's R points of group o are spaced by its lower span 2^lgl.  The bits of
// tl, c and r land on disjoint bits, so bpos(tl, c, r) = bpos(tl, 0, 0) +
// bpos(0, c, r).
template <class SH, int PH>
__device__ __forceinline__ int bpos(int tl, int c, int r) {
  constexpr int lgl = SH::lglam(SH::plo(PH));
  constexpr int lgR = SH::lgR(PH);
  const int o = group_of<SH, PH>(tl, c);
  return ((o >> lgl) << (lgl + lgR)) | (r << lgl) | (o & ((1 << lgl) - 1));
}

// the local index r of phase PH whose mixed digit reversal is k
template <class SH, int PH>
constexpr int rev_index(int k) {
  for (int r = 0; r < (1 << SH::lgR(PH)); ++r)
    if (SH::rev_phase(PH, r) == k) return r;
  return -1;
}

// Phase twiddle of phase PH.  The R points of a butterfly group take
// w^(mu * kr), kr = 0 .. R-1, with w = w_P^(2^sh); only the base b = w^mu
// comes from the two LDS tables of w_{2^14} (index e = (mu << sh) << (14 -
// log2 P); when the phase's stages span at most 128 points, e is a multiple
// of 128 and b = hi[e >> 7] exactly).  The other powers follow as products
// b^k = b^(k/2) * b^(k - k/2), at most four products deep: R - 2 complex
// multiplies instead of R - 1 table reads and products.
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
    static_assert(R <= 16, "phase radix");
    sfor<SH::E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      const int mu = group_of<SH, PH>(tl, c) & ((1 << lgl) - 1);
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
      // each power is applied as soon as it exists; only powers up to R / 2
      // feed later products, so at most R / 2 stay live
      sfor<R - 1>([&](auto k_) PACG_INL {
        constexpr int k = 1 + decltype(k_)::value;
        if constexpr (k >= 2) {
          constexpr int a = k / 2, b = k - k / 2;
          pr[k] = pr[a] * pr[b] - pi[a] * pi[b];
          pi[k] = pr[a] * pi[b] + pi[a] * pr[b];
        }
        constexpr int r = rev_index<SH, PH>(k);
        cmul<float>(vr[c * R + r], vi[c * R + r], pr[k], pi[k]);
      });
    });
  }
}

// Thread (within its convolution) that holds point s of a line in a phase
// whose lower stages span 2^lgl points, of radix 2^lgr, with 2^lgg groups per
// thread (group_of: the top phase deals groups out, the others keep them
// adjacent).
constexpr int holder(int s, int lgl, int lgr, int lgg, bool top, int tpl) {
  const int o = ((s >> (lgl + lgr)) << lgl) | (s & ((1 << lgl) - 1));
  return top ? o % tpl : o >> lgg;
}
// An exchange between phases A and B stays inside one wavefront when every
// point is held by threads of the same wave in both layouts (always when a
// convolution's threads fit in one wave).
template <class SH, int A, int B>
constexpr bool wave_local() {
  if (SH::TPL <= 64) return true;
  const int la = SH::lglam(SH::plo(A)), ra = SH::lgR(A), ga = SH::lgE() - ra;
  const int lb = SH::lglam(SH::plo(B)), rb = SH::lgR(B), gb = SH::lgE() - rb;
  const bool ta = A == SH::NP - 1, tb = B == SH::NP - 1;
  for (int s = 0; s < SH::P; ++s)
    if (holder(s, la, ra, ga, ta, SH::TPL) / 64 != holder(s, lb, rb, gb, tb, SH::TPL) / 64) return false;
  return true;
}

// A run of items of one shape SH starting at item i (the workgroup's list
// ends at `end`); returns the index of the first item not of this shape.
template <class SH>
__device__ __forceinline__ uint32_t batch_run(const pacg_batch_args_t& a, uint32_t i, const uint32_t end,
                                              float2* lds, const float2* twh, const float2* twl_, bool first) {
  constexpr int P = SH::P, E = SH::E, TPL = SH::TPL, NP = SH::NP;
  constexpr int G = BT / TPL;                // convolutions per item
  constexpr int IPTS = BT * E;               // points per item
  constexpr bool DBUF = IPTS <= BPTS / 2;    // two exchange buffers fit
  constexpr int PS = E == 16 ? 4 : 3;        // one pad slot per 2^PS points
  constexpr int BUFSZ = IPTS + (IPTS >> PS);
  static_assert(BT % TPL == 0 && (E == 8 || E == 16) && IPTS <= BPTS && (DBUF ? 2 : 1) * BUFSZ <= BLD,
                "item shape");
  // Per-thread indexes are re-derived (through an opaque copy of the thread
  // index) at every transform and exchange: derived once, every LDS slot and
  // global position of every phase would be loop-invariant, hoisted out of
  // the item loop and kept live, and the 16-point shapes would spill.
  int g, tl;
  auto fresh = [&]() PACG_INL {
    unsigned t = threadIdx.x;
    asm volatile("" : "+v"(t));
    g = (int)(t / (unsigned)TPL);
    tl = (int)(t % (unsigned)TPL);
  };
  fresh();
  const pacg_batch_item_t it0 = a.items[i];
  const uint32_t key = ((uint32_t)it0.log2n << 8) | it0.variant;
  const float scale = 1.0f / (float)P;
  const uint32_t bytes = a.total * 4u;
  const rsrc_t rkr = make_rsrc(a.kr, bytes), rki = make_rsrc(a.ki, bytes);
  const rsrc_t rxr = make_rsrc(a.xr, bytes), rxi = make_rsrc(a.xi, bytes);
  const rsrc_t ryr = make_rsrc(a.yr, bytes), ryi = make_rsrc(a.yi, bytes);

  // element offset of this thread's convolution in the item: its own pair, or
  // the item's first pair where a ragged item has fewer than G
  auto pair_base = [&](const pacg_batch_item_t& it) PACG_INL -> uint32_t {
    return a.conv_off[it.first + ((uint32_t)g < (uint32_t)it.count ? (uint32_t)g : 0u)];
  };
  // the points of the top (first DIF) phase from global memory, part Q of NQ
  // (the read-ahead is spread over the phases of a transform, so that the
  // workgroups do not all meet one burst of loads at each item boundary):
  // position pos(tl, 0, 0) + pos(0, c, r), the second part a constant
  auto load_top = [&](auto Q_, auto NQ_, rsrc_t br, rsrc_t bi, uint32_t b, float* dr, float* di) PACG_INL {
    constexpr int PH = NP - 1;
    constexpr int R = 1 << SH::lgR(PH);
    constexpr int Q = decltype(Q_)::value, NQ = decltype(NQ_)::value;
    constexpr int E0 = Q * E / NQ, E1 = (Q + 1) * E / NQ;
    const uint32_t vo = (b + (uint32_t)bpos<SH, PH>(tl, 0, 0)) * 4u;
    sfor<E1 - E0>([&](auto j_) PACG_INL {
      constexpr int e = E0 + decltype(j_)::value;
      constexpr int c = e / R, r = e % R;
      const uint32_t so = (uint32_t)bpos<SH, PH>(0, c, r) * 4u;  // folds to a constant
      dr[e] = buf_load(br, vo, so);
      di[e] = buf_load(bi, vo, so);
    });
  };
  using One = std::integral_constant<int, 1>;
  using Zero = std::integral_constant<int, 0>;
  // the read-ahead of a transform is issued in NQ parts at its phases 0 ..
  // NQ - 1: the last part still has a whole phase to arrive
  constexpr int NQ = NP > 1 ? NP - 1 : 1;
  using NQc = std::integral_constant<int, NQ>;

  float vr[E], vi[E];  // working points
  // exchange number X of the item: through LDS buffer X % 2 when two fit (the
  // buffer it writes was last read two exchanges ago, before the barrier of
  // the one in between: one barrier), else through the one buffer (two).
  // With one buffer, an exchange that stays inside each wavefront needs no
  // workgroup barrier: the slots a wave writes were last read by that same
  // wave (in the layout this exchange starts from), and its own LDS accesses
  // complete in order once it waits for its writes.
  auto exchange = [&](auto FROM_, auto TO_, auto X_) PACG_INL {
    constexpr int FROM = decltype(FROM_)::value;
    constexpr int TO = decltype(TO_)::value;
    constexpr int RF = 1 << SH::lgR(FROM);
    constexpr int RT = 1 << SH::lgR(TO);
    constexpr int BUF = DBUF ? (decltype(X_)::value & 1) * BUFSZ : 0;
    constexpr bool LOCAL = !DBUF && wave_local<SH, FROM, TO>();
    if constexpr (!DBUF && !LOCAL) __syncthreads();
    fresh();
    {
      float2* w = lds + BUF + lds_pad<PS>(g * P + bpos<SH, FROM>(tl, 0, 0));
      sfor<E / RF>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<RF>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          w[lds_pad<PS>(bpos<SH, FROM>(0, c, r))] = make_float2(vr[c * RF + r], vi[c * RF + r]);
        });
      });
    }
    if constexpr (LOCAL)
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // this wave's writes done; no reordering
    else
      __syncthreads();
    fresh();
    {
      const float2* w = lds + BUF + lds_pad<PS>(g * P + bpos<SH, TO>(tl, 0, 0));
      sfor<E / RT>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<RT>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const float2 v = w[lds_pad<PS>(bpos<SH, TO>(0, c, r))];
          vr[c * RT + r] = v.x;
          vi[c * RT + r] = v.y;
        });
      });
    }
  };
  // A^T of the working points; X0: number of the first exchange; hook(i) runs
  // at the start of the i-th phase
  auto dif = [&](auto X0_, auto&& hook) PACG_INL {
    fresh();
    sfor<NP>([&](auto i_) PACG_INL {
      constexpr int PH = NP - 1 - decltype(i_)::value;
      hook(i_);
      if constexpr (PH < NP - 1)
        exchange(std::integral_constant<int, PH + 1>{}, std::integral_constant<int, PH>{},
                 std::integral_constant<int, decltype(X0_)::value + NP - 2 - PH>{});
      phase_stages<float, SH, PH, M_DIF>(vr, vi);
      batch_phase_twiddle<SH, PH, false>(vr, vi, tl, twh, twl_);
    });
  };

  float kr[E], ki[E];  // k of the current item, loaded ahead
  pacg_batch_item_t it = it0;
  uint32_t base = pair_base(it);
  load_top(Zero{}, One{}, rkr, rki, base, kr, ki);
  if (first) __syncthreads();  // the twiddle tables, written at launch, are visible (k already in flight)
  for (;;) {
    // The next item's record is read now and its pair offset after A^T k:
    // each has a whole transform to arrive before it is used.
    const uint32_t nx = i + 1;
    const bool has_next = nx < end;
    uint2 nrec = make_uint2(0u, 0u);
    if (has_next) nrec = reinterpret_cast<const uint2*>(a.items)[nx];
    fresh();
    float xr[E], xi[E];  // x arrives while k is transformed, a part per phase
#pragma unroll
    for (int e = 0; e < E; ++e) {
      vr[e] = kr[e];
      vi[e] = ki[e];
    }
    dif(Zero{}, [&](auto i_) PACG_INL {
      if constexpr (decltype(i_)::value < NQ) load_top(i_, NQc{}, rxr, rxi, base, xr, xi);
    });  // A^T k
    // the record is the same in every lane: make it scalar, so that the
    // branches on it and the exit of this loop are uniform
    pacg_batch_item_t itn;
    {
      const uint32_t lo = __builtin_amdgcn_readfirstlane(nrec.x), hi = __builtin_amdgcn_readfirstlane(nrec.y);
      itn.first = lo;
      itn.count = (uint16_t)(hi & 0xFFFFu);
      itn.log2n = (uint8_t)((hi >> 16) & 0xFFu);
      itn.variant = (uint8_t)(hi >> 24);
    }
    const bool more = has_next && (((uint32_t)itn.log2n << 8) | itn.variant) == key;
    uint32_t nbase = base;
    if (more) nbase = pair_base(itn);
    float sr[E], si[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      sr[e] = vr[e] * scale;
      si[e] = vi[e] * scale;
      vr[e] = xr[e];
      vi[e] = xi[e];
    }
    dif(std::integral_constant<int, NP - 1>{}, [&](auto) PACG_INL {});  // A^T x
#pragma unroll
    for (int e = 0; e < E; ++e) cmul<float>(vr[e], vi[e], sr[e], si[e]);
    // conj(A); the next item's k arrives during it, a part per phase
    fresh();
    sfor<NP>([&](auto i_) PACG_INL {
      constexpr int PH = decltype(i_)::value;
      if constexpr (PH < NQ) {
        if (more) load_top(i_, NQc{}, rkr, rki, nbase, kr, ki);
      }
      if constexpr (PH > 0)
        exchange(std::integral_constant<int, PH - 1>{}, std::integral_constant<int, PH>{},
                 std::integral_constant<int, 2 * (NP - 1) + PH - 1>{});
      batch_phase_twiddle<SH, PH, true>(vr, vi, tl, twh, twl_);
      phase_stages<float, SH, PH, M_DITC>(vr, vi);
    });
    {
      constexpr int PH = NP - 1;
      constexpr int R = 1 << SH::lgR(PH);
      fresh();
      const bool valid = (uint32_t)g < (uint32_t)it.count;
      const uint32_t vo = (base + (uint32_t)bpos<SH, PH>(tl, 0, 0)) * 4u;
      sfor<E / R>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<R>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const uint32_t so = (uint32_t)bpos<SH, PH>(0, c, r) * 4u;  // folds to a constant
          if (valid) {
            buf_store(vr[c * R + r], ryr, vo, so);
            buf_store(vi[c * R + r], ryi, vo, so);
          }
        });
      });
    }
    __syncthreads();  // LDS is reused by the next item
    if (!more) return nx;
    i = nx;
    it = itn;
    base = nbase;
  }
}

// Variant table: (log2 n, variant) -> shape (stage count, 2-bit log radices
// with stage 0 first, phase count, 4-bit top stage of each phase with phase 0
// first).  Variant 0 of every length holds 16 points per thread (phases of
// radix 4 x 4, one phase fewer than 8-point phases: fewer exchanges and
// fewer phase twiddles); at 2^13 and 2^14 its phases below the top span 1024
// points, one wavefront's share, so only the exchange with the top phase
// needs the workgroup barrier.  The other variants are alternatives for the
// autotuner (8 points per thread where listed).  The last two columns are
// the time of one item on one workgroup in 0.1 us on MI355X: measured with
// batches of one length (2^24 points), the costs of the default dealing; and
// fitted (least squares, with a cost per run of one shape) to per-workgroup
// timings of batches of mixed lengths 2^U{8..14}, where short items run beside
// long ones and cost relatively less, the costs of the slice dealing.  Kept in
// step with pacg_batch_variants[] below.
#define PACG_BATCH_VARIANTS(X)                                                \
  X(8, 0, 4, 0xAAu, 2, 0x31u, 201, 135)         /* 4 4 | 4 4                */  \
  X(8, 1, 3, 0x2Fu, 3, 0x210u, 120, 54)         /* 8 | 8 | 4                */  \
  X(9, 0, 5, 0x1AAu, 3, 0x431u, 197, 137)       /* 4 4 | 4 4 | 2            */  \
  X(9, 1, 3, 0x3Fu, 3, 0x210u, 124, 71)         /* 8 | 8 | 8                */  \
  X(10, 0, 5, 0x2AAu, 3, 0x431u, 195, 137)      /* 4 4 | 4 4 | 4            */  \
  X(10, 1, 4, 0xAFu, 4, 0x3210u, 147, 69)       /* 8 | 8 | 4 | 4            */  \
  X(11, 0, 5, 0x3AAu, 3, 0x431u, 249, 169)      /* 4 4 | 4 4 | 8            */  \
  X(11, 1, 4, 0xBFu, 4, 0x3210u, 143, 121)      /* 8 | 8 | 8 | 4            */  \
  X(12, 0, 6, 0xAAAu, 3, 0x531u, 251, 171)      /* 4 4 | 4 4 | 4 4          */  \
  X(12, 1, 4, 0xFFu, 4, 0x3210u, 144, 114)      /* 8 | 8 | 8 | 8            */  \
  X(13, 0, 6, 0xEAAu, 4, 0x5431u, 298, 228)     /* 4 4 | 4 4 | 4 | 8        */  \
  X(13, 1, 7, 0x1AAAu, 4, 0x6531u, 305, 218)    /* 4 4 | 4 4 | 4 4 | 2      */  \
  X(13, 2, 5, 0x1FFu, 5, 0x43210u, 182, 127)    /* 8 | 8 | 8 | 8 | 2        */  \
  X(14, 0, 7, 0x2AAAu, 4, 0x6431u, 290, 205)    /* 4 4 | 4 4 | 4 | 4 4      */  \
  X(14, 1, 7, 0x2AAAu, 4, 0x6531u, 305, 228)    /* 4 4 | 4 4 | 4 4 | 4      */  \
  X(14, 2, 7, 0x2AAAu, 4, 0x6420u, 288, 208)    /* 4 | 4 4 | 4 4 | 4 4      */

__global__ void __launch_bounds__(BT) batch_kernel(const pacg_batch_args_t a) {
  __shared__ float2 lds[BLD];
  __shared__ float2 twh[128], twl[128];
  {
    const float2* tw = reinterpret_cast<const float2*>(a.tw);
    const int t = threadIdx.x;
    if (t < 128) twh[t] = tw[t * 128];
    else if (t < 256) twl[t - 128] = tw[t - 128];
  }
  uint32_t i = a.wg_first[blockIdx.x];
  const uint32_t end = a.wg_first[blockIdx.x + 1];
  if (i >= end) return;
  bool first = true;  // the first run waits for the tables after issuing its first loads
  while (i < end) {
    const pacg_batch_item_t it = a.items[i];
    const int key = (int)it.log2n * 4 + (int)it.variant;
    switch (key) {
#define PACG_CASE(M, V, NS, SC, NPH, PC, COST, MIXED)                   \
  case M * 4 + V:                                                       \
    static_assert(Shape<NS, SC, NPH, PC>::LGP == M, "variant length"); \
    i = batch_run<Shape<NS, SC, NPH, PC>>(a, i, end, lds, twh, twl, first); \
    break;
      PACG_BATCH_VARIANTS(PACG_CASE)
#undef PACG_CASE
      default:
        i = end;  // no such variant (the host never builds one): leave the list
        break;
    }
    first = false;
  }
}

}  // namespace pacg

extern "C" {

// Stage schedules of the variants (stage 0 first), for the host and tests.
#define PACG_SH(NS, SC, NPH, PC) pacg::Shape<NS, SC, NPH, PC>
#define PACG_RHO(NS, SC, NPH, PC, G) (NS > G ? PACG_SH(NS, SC, NPH, PC)::rho(G) : 0)
#define PACG_ROW(M, V, NS, SC, NPH, PC, COST, MIXED)                                                     \
  {M, V, NS,                                                                                             \
   {PACG_RHO(NS, SC, NPH, PC, 0), PACG_RHO(NS, SC, NPH, PC, 1), PACG_RHO(NS, SC, NPH, PC, 2),            \
    PACG_RHO(NS, SC, NPH, PC, 3), PACG_RHO(NS, SC, NPH, PC, 4), PACG_RHO(NS, SC, NPH, PC, 5),            \
    PACG_RHO(NS, SC, NPH, PC, 6)},                                                                       \
   PACG_SH(NS, SC, NPH, PC)::TPL, COST, MIXED},
const pacg_batch_variant_t pacg_batch_variants[] = {PACG_BATCH_VARIANTS(PACG_ROW)};
#undef PACG_RHO
#undef PACG_SH
#undef PACG_ROW
const int pacg_num_batch_variants = sizeof(pacg_batch_variants) / sizeof(pacg_batch_variants[0]);

int pacg_batch_launch(const pacg_batch_args_t* a, int blocks, void* stream) {
  if (a->n_items == 0 || blocks < 1) return 0;
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
