// Multi-stage butterfly pass kernels for gfx950.
//
// A transform along one axis is a product of radix stages (reference
// butterflies.c): stage q with radix r_q, l_q = prod_{s<q} r_s, k_q = r_q l_q
// maps legs m of the block b at positions b k + m l + j by
//   A^T (DIF, butterflies.c:70-97,217-275,505-632):  y = F_r x,  y_m *= w_k^{m j}
//   A   (DIT, butterflies.c:12-39,99-156,277-389):   x_m *= w_k^{m j}, y = F_r x
//   conj(A)   (butterflies.c:41-68,158-215,391-503): conjugate of the above.
//
// A *pass* executes a run of consecutive stages [q_lo, q_hi] of one axis.
// Its P = prod r_q points of one line sit S elements apart; the pass reads
// every line once, does all its stages in registers and LDS, and writes the
// line back to the same positions.  Two identities keep the twiddles cheap
// (both exact in real arithmetic):
//   * the stages of a pass equal a P-point DIF on the line followed by the
//     diagonal w_K^{j' k(s)}, K = P l_{q_lo}, j' the within-axis position below
//     the pass and k(s) the mixed digit reversal of the point index s;
//   * the same holds one level down: a register *phase* (stages one thread
//     holds together) is an R-point DIF with compile-time twiddles followed by
//     w_{R lambda}^{mu k(r)} with mu the thread's lower sub-index.
// DIT passes are the transposes: diagonal first, then the stages bottom up.
// The fused pass of the convolution runs A^T, multiplies by the spectrum at
// the same positions and runs conj(A) -- positions never move, so no
// digit-reversal permutation is ever executed.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <utility>

#include "pacg_internal.h"

// every lambda of the pass code is inlined: the register arrays must never escape
#define PACG_INL __attribute__((always_inline))

namespace pacg {

// ---------------------------------------------------------------------------
// compile-time roots of unity: cos/sin(2 pi x / 64), octant-reduced Taylor sums
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;

constexpr double sin_poly(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k <= 13; ++k) {
    term = -term * x2 / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cos_poly(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k <= 13; ++k) {
    term = -term * x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
constexpr double cs64(int x, bool want_sin) {
  x &= 63;
  int q = x >> 4, rem = x & 15;
  double c = 0.0, s = 0.0;
  if (rem == 0) {
    c = 1.0;
    s = 0.0;
  } else if (rem == 8) {
    c = 0.70710678118654752440;
    s = 0.70710678118654752440;
  } else if (rem < 8) {
    c = cos_poly(rem * kPi / 32.0);
    s = sin_poly(rem * kPi / 32.0);
  } else {
    c = sin_poly((16 - rem) * kPi / 32.0);
    s = cos_poly((16 - rem) * kPi / 32.0);
  }
  for (int i = 0; i < q; ++i) {
    double t = c;
    c = -s;
    s = t;
  }
  return want_sin ? s : c;
}

// ---------------------------------------------------------------------------
// compile-time loops
// ---------------------------------------------------------------------------
template <typename F, int... I>
__device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// pass shape: NS stages (2 bits each: log2 radix, stage 0 = unit sub-stride),
// NP register phases (4 bits each: index of the phase's top stage);
// LGE raises the points per thread above the largest phase when two passes
// share one thread block (pair kernels)
// ---------------------------------------------------------------------------
template <int NS_, uint32_t SC, int NP_, uint32_t PC, int LGE = 0>
struct Shape {
  static constexpr int NS = NS_;
  static constexpr int NP = NP_;
  static constexpr int lg(int g) { return (int)((SC >> (2 * g)) & 3u); }
  static constexpr int rho(int g) { return 1 << lg(g); }
  static constexpr int lglam(int g) {
    int v = 0;
    for (int h = 0; h < g; ++h) v += lg(h);
    return v;
  }
  static constexpr int LGP = lglam(NS);
  static constexpr int P = 1 << LGP;
  static constexpr int phi(int p) { return (int)((PC >> (4 * p)) & 15u); }
  static constexpr int plo(int p) { return p == 0 ? 0 : phi(p - 1) + 1; }
  static constexpr int lgR(int p) { return lglam(phi(p) + 1) - lglam(plo(p)); }
  static constexpr int lgE() {
    int m = LGE;
    for (int p = 0; p < NP; ++p) m = lgR(p) > m ? lgR(p) : m;
    return m;
  }
  static constexpr int E = 1 << lgE();
  static constexpr int TPL = P / E;
  // digit reversal of the local index r within phase p (mixed radix)
  static constexpr int rev_phase(int p, int r) {
    int lo = plo(p), out = 0;
    int R = 1 << lgR(p);
    for (int g = lo; g <= phi(p); ++g) {
      int sh = lglam(g) - lglam(lo);
      int d = (r >> sh) & (rho(g) - 1);
      int kk = 1 << (lglam(g + 1) - lglam(lo));
      out += d * (R / kk);
    }
    return out;
  }
  static constexpr bool valid() {
    if (NS < 1 || NS > 12 || NP < 1 || phi(NP - 1) != NS - 1) return false;
    for (int p = 0; p < NP; ++p)
      if (phi(p) < plo(p)) return false;
    for (int g = 0; g < NS; ++g)
      if (lg(g) == 0) return false;
    return E <= 16 && E <= P && P <= 16384;
  }
};

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cmul(T& ar, T& ai, T br, T bi) {
  T r = ar * br - ai * bi;
  T i = ar * bi + ai * br;
  ar = r;
  ai = i;
}

// multiply v by exp(-2 pi i x / 64), conjugated when CJ; exact for quarter turns
template <typename T, int X, bool CJ>
__device__ __forceinline__ void rot64(T& vr, T& vi) {
  constexpr int x = X & 63;
  if constexpr (x == 0) {
    return;
  } else if constexpr (x == 16) {  // -i (or +i)
    T r = vr;
    if constexpr (!CJ) {
      vr = vi;
      vi = -r;
    } else {
      vr = -vi;
      vi = r;
    }
  } else if constexpr (x == 32) {
    vr = -vr;
    vi = -vi;
  } else if constexpr (x == 48) {  // +i (or -i)
    T r = vr;
    if constexpr (!CJ) {
      vr = -vi;
      vi = r;
    } else {
      vr = vi;
      vi = -r;
    }
  } else {
    constexpr double c = cs64(x, false);
    constexpr double s = CJ ? cs64(x, true) : -cs64(x, true);
    cmul<T>(vr, vi, (T)c, (T)s);
  }
}

// In-place R-point DFT (R in 2,4,8) on v[B], v[B+ST], ..., natural order;
// INV selects exp(+2 pi i / R) (conjugate butterflies).
template <typename T, int R, bool INV, int B, int ST>
__device__ __forceinline__ void dft(T* vr, T* vi) {
  if constexpr (R == 2) {
    T ar = vr[B], ai = vi[B], br = vr[B + ST], bi = vi[B + ST];
    vr[B] = ar + br;
    vi[B] = ai + bi;
    vr[B + ST] = ar - br;
    vi[B + ST] = ai - bi;
  } else if constexpr (R == 4) {
    T t0r = vr[B] + vr[B + 2 * ST], t0i = vi[B] + vi[B + 2 * ST];
    T t1r = vr[B] - vr[B + 2 * ST], t1i = vi[B] - vi[B + 2 * ST];
    T t2r = vr[B + ST] + vr[B + 3 * ST], t2i = vi[B + ST] + vi[B + 3 * ST];
    T t3r = vr[B + ST] - vr[B + 3 * ST], t3i = vi[B + ST] - vi[B + 3 * ST];
    vr[B] = t0r + t2r;
    vi[B] = t0i + t2i;
    vr[B + 2 * ST] = t0r - t2r;
    vi[B + 2 * ST] = t0i - t2i;
    // forward: y1 = t1 - i t3, y3 = t1 + i t3
    if constexpr (!INV) {
      vr[B + ST] = t1r + t3i;
      vi[B + ST] = t1i - t3r;
      vr[B + 3 * ST] = t1r - t3i;
      vi[B + 3 * ST] = t1i + t3r;
    } else {
      vr[B + ST] = t1r - t3i;
      vi[B + ST] = t1i + t3r;
      vr[B + 3 * ST] = t1r + t3i;
      vi[B + 3 * ST] = t1i - t3r;
    }
  } else {
    static_assert(R == 8, "radix");
    // even / odd halves: E = F4(x0,x2,x4,x6), O = F4(x1,x3,x5,x7)
    dft<T, 4, INV, B, 2 * ST>(vr, vi);
    dft<T, 4, INV, B + ST, 2 * ST>(vr, vi);
    // y_p = E_p + w8^p O_p, y_{p+4} = E_p - w8^p O_p ; E_p at B+2p*ST, O_p at B+(2p+1)*ST
    T er[4], ei[4], orr[4], oi[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      er[p] = vr[B + 2 * p * ST];
      ei[p] = vi[B + 2 * p * ST];
      orr[p] = vr[B + (2 * p + 1) * ST];
      oi[p] = vi[B + (2 * p + 1) * ST];
    }
    rot64<T, 8, INV>(orr[1], oi[1]);
    rot64<T, 16, INV>(orr[2], oi[2]);
    rot64<T, 24, INV>(orr[3], oi[3]);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      vr[B + p * ST] = er[p] + orr[p];
      vi[B + p * ST] = ei[p] + oi[p];
      vr[B + (p + 4) * ST] = er[p] - orr[p];
      vi[B + (p + 4) * ST] = ei[p] - oi[p];
    }
  }
}

enum { M_DIF = 0, M_DIT = 1, M_DITC = 2 };

// One stage G of phase PH applied to every group the thread holds.
template <typename T, class SH, int PH, int G, int MODE>
__device__ __forceinline__ void stage(T* vr, T* vi) {
  constexpr int lgR = SH::lgR(PH);
  constexpr int R = 1 << lgR;
  constexpr int lo = SH::plo(PH);
  constexpr int lgl = SH::lglam(G) - SH::lglam(lo);
  constexpr int rho = SH::rho(G);
  constexpr int NGRP = SH::E / R;
  constexpr int ST = 1 << lgl;
  constexpr int M = rho << lgl;  // modulus of the internal twiddle
  sfor<NGRP>([&](auto C_) PACG_INL {
    sfor<R / rho>([&](auto B_) PACG_INL {
      constexpr int bidx = decltype(B_)::value;
      constexpr int mu = bidx & (ST - 1);
      constexpr int hi = bidx >> lgl;
      constexpr int base = decltype(C_)::value * R + hi * M + mu;
      if constexpr (MODE != M_DIF) {
        sfor<rho>([&](auto m_) PACG_INL {
          constexpr int m = decltype(m_)::value;
          rot64<T, (m * mu) * (64 / M), MODE == M_DITC>(vr[base + m * ST], vi[base + m * ST]);
        });
      }
      dft<T, rho, MODE == M_DITC, base, ST>(vr, vi);
      if constexpr (MODE == M_DIF) {
        sfor<rho>([&](auto m_) PACG_INL {
          constexpr int m = decltype(m_)::value;
          rot64<T, (m * mu) * (64 / M), false>(vr[base + m * ST], vi[base + m * ST]);
        });
      }
    });
  });
}

// All stages of phase PH: DIF top-down, DIT bottom-up.
template <typename T, class SH, int PH, int MODE>
__device__ __forceinline__ void phase_stages(T* vr, T* vi) {
  constexpr int lo = SH::plo(PH), hi = SH::phi(PH);
  constexpr int cnt = hi - lo + 1;
  sfor<cnt>([&](auto i_) PACG_INL {
    constexpr int i = decltype(i_)::value;
    constexpr int g = (MODE == M_DIF) ? (hi - i) : (lo + i);
    stage<T, SH, PH, g, MODE>(vr, vi);
  });
}

// Phase-external twiddle w_{R lambda}^{mu k(r)}; tw(idx, wr, wi) returns w_P^idx.
// SERIAL 1: finish one group before the next group's table reads are issued;
// 2: one twiddle at a time (keeps register pressure down where each twiddle
// takes two table reads and a product, as for FP64 lines of 16-point phases).
template <typename T, class SH, int PH, bool CJ, int SERIAL = 0, typename TW>
__device__ __forceinline__ void phase_twiddle_f(T* vr, T* vi, int tl, TW&& tw) {
  constexpr int lo = SH::plo(PH);
  if constexpr (lo > 0) {
    constexpr int lgR = SH::lgR(PH);
    constexpr int R = 1 << lgR;
    constexpr int lgl = SH::lglam(lo);
    constexpr int sh = SH::LGP - SH::lglam(SH::phi(PH) + 1);
    sfor<SH::E / R>([&](auto C_) PACG_INL {
      constexpr int c = decltype(C_)::value;
      const int o = tl + SH::TPL * c;
      const int mu = o & ((1 << lgl) - 1);
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        constexpr int kr = SH::rev_phase(PH, r);
        if constexpr (kr != 0) {
          const int idx = (mu * kr) << sh;
          T wr, wi;
          tw(idx, wr, wi);
          if constexpr (CJ) wi = -wi;
          cmul<T>(vr[c * R + r], vi[c * R + r], wr, wi);
          if constexpr (SERIAL == 2) __builtin_amdgcn_sched_barrier(0);
        }
      });
      if constexpr (SERIAL == 1) __builtin_amdgcn_sched_barrier(0);
    });
  }
}

// ... from the w_P table in LDS
template <typename T, class SH, int PH, bool CJ>
__device__ __forceinline__ void phase_twiddle(T* vr, T* vi, const T* twr, const T* twi, int tl) {
  phase_twiddle_f<T, SH, PH, CJ>(vr, vi, tl, [&](int idx, T& wr, T& wi) PACG_INL {
    wr = twr[idx];
    wi = twi[idx];
  });
}

// Position (within the line) of element (c, r) of thread tl in phase PH.
template <class SH, int PH>
__device__ __forceinline__ int pos(int tl, int c, int r) {
  constexpr int lgl = SH::lglam(SH::plo(PH));
  constexpr int lgR = SH::lgR(PH);
  const int o = tl + SH::TPL * c;
  const int mu = o & ((1 << lgl) - 1);
  const int beta = o >> lgl;
  return (beta << (lgl + lgR)) | (r << lgl) | mu;
}

// LDS slot of point s in an exchange tile without padding: bits 0-4 of s are
// XOR-ed with h(bits 5-7), h linear over GF(2) with h(1) = 31, h(2) = 21,
// h(4) = 25.  For every phase layout of radix-2/4/8 schedules (pos above) the
// 32 lanes of a half-wave then reach 32 different banks (32 banks of 4 bytes,
// a wave64 access served as two half-waves).  Only bits 0-4 change, so the map
// is a bijection on every aligned block of 32 points; it is linear, so the slot
// of pos(tl, c, r) = pos(tl, 0, 0) | pos(0, c, r) (disjoint bits) is the XOR of
// a per-thread part and a compile-time constant.
__device__ __forceinline__ int lds_swz(int s) {
  const int h = (-((s >> 5) & 1) & 31) ^ (-((s >> 6) & 1) & 21) ^ (-((s >> 7) & 1) & 25);
  return s ^ h;
}

// mixed digit reversal of a point index s over all stages of the pass
template <class SH>
__device__ __forceinline__ int rev_pass(int s) {
  int out = 0;
  sfor<SH::NS>([&](auto g_) PACG_INL {
    constexpr int g = decltype(g_)::value;
    const int d = (s >> SH::lglam(g)) & (SH::rho(g) - 1);
    out |= d << (SH::LGP - SH::lglam(g + 1));
  });
  return out;
}

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nb) {
  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD one
  // contiguous range of tiles so neighbouring tiles share its L2.
  if ((nb & 7u) != 0u) return bid;
  return (bid & 7u) * (nb >> 3) + (bid >> 3);
}

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

// 16-byte vectors: 4 floats or 2 doubles
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
  static __device__ __forceinline__ void split(const float4& v, float* o) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
  static __device__ __forceinline__ float4 join(const float* o) { return make_float4(o[0], o[1], o[2], o[3]); }
  // a 16-byte non-temporal load (the builtin takes native vector types)
  static __device__ __forceinline__ float4 load_nt(const float* p) {
    typedef float nv __attribute__((ext_vector_type(4)));
    const nv t = __builtin_nontemporal_load(reinterpret_cast<const nv*>(p));
    return make_float4(t[0], t[1], t[2], t[3]);
  }
};
template <>
struct Vec4<double> {
  using type = double2;
  static __device__ __forceinline__ void split(const double2& v, double* o) {
    o[0] = v.x;
    o[1] = v.y;
  }
  static __device__ __forceinline__ double2 join(const double* o) { return make_double2(o[0], o[1]); }
  static __device__ __forceinline__ double2 load_nt(const double* p) {
    typedef double nv __attribute__((ext_vector_type(2)));
    const nv t = __builtin_nontemporal_load(reinterpret_cast<const nv*>(p));
    return make_double2(t[0], t[1]);
  }
};

// Row bits of a strided FP32 tile with 16 lines (one bit) or 8 lines (two
// bits) that are XOR-ed into the row of point s, so that the 2 or 4 rows a
// half-wave touches in any radix-2/4/8 phase layout fall in different banks:
// 16 lines, row bit 0 ^= parity(s bits 2-4); 8 lines, row bit 0 ^= parity(s
// bits 2, 4) and row bit 1 ^= parity(s bits 3, 5).  Linear over GF(2).
template <int LINES>
__device__ __forceinline__ int row_flip(int s) {
  if constexpr (LINES == 16)
    return __builtin_parity((unsigned)s & 0x1Cu);
  else
    return __builtin_parity((unsigned)s & 0x14u) | (__builtin_parity((unsigned)s & 0x28u) << 1);
}

enum { LM_PAD = 0, LM_ROWS = 1, LM_ROWSWZ = 2, LM_LINESWZ = 3, LM_ROWXOR = 4, LM_QUADSWZ = 5 };

// Slot of point s in a contiguous FP64 line of (re, im) pairs (16-byte slots,
// eight to 32 banks of 4 bytes) whose phases below the top are single radix-4
// stages: bits 0-3 of s become s0^s4, s1^s3, s2^s4, s3^s5.  Consecutive
// threads then hold points 4 apart (phase 0), in runs of 4 that are 16 apart
// (phase 1) or in runs of 16 or more (the others), and in every one of these
// layouts each aligned group of 8 or 16 lanes meets 8 or 16 different slots
// modulo 8 or 16, where one pad slot per 8 points leaves 2-way conflicts.
// Each bit is XOR-ed only with bits above it, so the map is a bijection; it is
// linear, so the slot of pos(tl, 0, 0) | pos(0, c, r) is the XOR of their slots.
__device__ __forceinline__ constexpr int quad_swz(int s) {
  return s ^ (((s >> 4) & 1) | (((s >> 3) & 1) << 1) | (((s >> 4) & 1) << 2) | (((s >> 5) & 1) << 3));
}
template <class SH>
constexpr bool radix4_phases() {
  for (int p = 0; p + 1 < SH::NP; ++p)
    if (SH::plo(p) != SH::phi(p) || SH::lg(SH::plo(p)) != 2) return false;
  return SH::NP > 1;
}

// Row of point s in a strided FP64 tile of long lines ([P][LINES], no
// padding): the low NB bits of s are XOR-ed with every NB-bit group of s above
// them.  In the bottom phases consecutive threads hold points R rows apart;
// the fold spreads them over distinct rows modulo 2^NB, and for every phase
// layout of the 2048-point shapes the 32 lanes of a half-wave meet each of
// the 32 banks equally often (NB = 3 up to 4 lines, 4 for 8 lines).  Linear
// over GF(2), like the other layouts.
template <int NB, int LGP>
__device__ __forceinline__ constexpr int row_fold(int s) {
  int h = 0;
  for (int k = NB; k < LGP; k += NB) h ^= s >> k;  // a fixed count: no loop at run time
  return s ^ (h & ((1 << NB) - 1));
}

// ---------------------------------------------------------------------------
// the pass kernel
// ---------------------------------------------------------------------------
// Minimum waves per SIMD the register allocation must allow.  FP64 tiles of
// lines of 2048+ points: LDS leaves room for 4 waves per SIMD (strided: two
// 512-thread or four 256-thread workgroups per CU; contiguous: four one-line
// workgroups), and without the bound the allocator takes up to 254 VGPRs for
// the smaller strided workgroups (2 waves per SIMD) and 130 for the fused
// contiguous one of 8 points per thread (3 waves).  Other tiles: no bound.
template <typename T, class SH, int LAYOUT>
constexpr int pass_min_waves() {
  return (sizeof(T) == 8 && SH::P >= 2048) ? 4 : 1;
}

// NTM, the load policy (chosen at launch from the plan's flags; a separate
// instantiation, since a branch on a flag inside one kernel costs registers in
// both forms): bit 0, the fused pass streams the spectrum with non-temporal
// loads; bit 1, an A^T pass loads its input non-temporally (the caller's x,
// which one convolution reads once); bit 3, NO_EXT, the pass applies no
// pass-external twiddle (a.ext == 0) and its kernel leaves that code out: the
// FP64 strided conj(A) and A tiles of 2048-point lines, which start with it,
// need 105 instead of 128 VGPRs without it and no longer spill (26 VGPRs).
enum { NT_SPEC = 1, NT_IN = 2, NO_EXT = 8 };
template <typename T, typename TS, class SH, int KIND, int LAYOUT, int LINES, int NTM = 0>
__global__ void __launch_bounds__(LINES* SH::TPL) __attribute__((amdgpu_waves_per_eu(pass_min_waves<T, SH, LAYOUT>())))
    pass_kernel(const pacg_pass_args_t a) {
  static_assert(SH::valid(), "invalid pass shape");
  constexpr int P = SH::P, E = SH::E, TPL = SH::TPL, NP = SH::NP;
  constexpr int NT = LINES * TPL;
  // LDS exchange tiles.  FP32 tiles are laid out so that no exchange of a
  // radix-2/4/8 phase layout meets a bank conflict (32 banks of 4 bytes, a
  // wave64 access served as two half-waves of 32 lanes):
  //   LM_ROWS    strided, 32 lines or more: [P][LINES], a half-wave reads one row;
  //   LM_ROWSWZ  strided, 8 or 16 lines: [P][LINES], the low log2(32 / LINES)
  //              bits of the row XOR-ed with parities of the point index (row_flip);
  //   LM_LINESWZ contiguous, 256 points or more at most 8 per thread, or 2048
  //              points: [LINES][P] with lds_swz inside each line.
  // FP64 strided tiles of lines of 2048+ points (a few lines wide, so that a
  // tile fits in LDS at all) use LM_ROWXOR: [P][LINES] with the row folded by
  // row_fold, no padding.  FP64 contiguous lines of 2048+ points, 8 points
  // per thread, in single radix-4 phases below the top use LM_QUADSWZ:
  // [LINES][P] with quad_swz inside each line.  Other tiles keep LM_PAD:
  // strided [P][LINES+1], contiguous [LINES][P + P/8 + 1] with one pad slot
  // per 8 points.
  constexpr bool F32 = sizeof(T) == 4;
  constexpr bool LONG_S = !F32 && LAYOUT == PACG_L_STRIDED && P >= 2048;
  constexpr int LMODE = LONG_S ? LM_ROWXOR
                        : !F32  ? ((LAYOUT == PACG_L_CONTIG && P >= 2048 && E == 8 && radix4_phases<SH>()) ? LM_QUADSWZ
                                                                                                          : LM_PAD)
                        : LAYOUT == PACG_L_STRIDED ? (LINES >= 32                  ? LM_ROWS
                                                      : (LINES == 16 || LINES == 8) ? LM_ROWSWZ
                                                                                    : LM_PAD)
                        : ((E <= 8 && P >= 256) || P >= 2048) ? LM_LINESWZ
                                                              : LM_PAD;
  constexpr int XNB = LINES >= 8 ? 4 : 3;  // fold width of LM_ROWXOR
  constexpr int LDJ = (LMODE == LM_ROWS || LMODE == LM_ROWSWZ || LMODE == LM_ROWXOR) ? LINES : LINES + 1;
  constexpr int LDP = (LMODE == LM_LINESWZ || LMODE == LM_QUADSWZ) ? P : P + P / 8 + 1;
  constexpr int LDS_N = (LAYOUT == PACG_L_STRIDED) ? P * LDJ : LINES * LDP;
  constexpr bool NEED_X = NP > 1;
  // FP64 contiguous lines of 2048+ points with 16 points per thread exchange
  // the real and the imaginary parts one after the other through one LDS
  // array: half the LDS per line, so twice the workgroups per CU (2 -> 4 waves
  // per SIMD, where LDS and not registers bounded the occupancy; 44.5 -> 37.6
  // us per pass at 2048^2).  With 8 points per thread (the default for these
  // lines) the paired exchange below is faster (about 31 against 33 us), and
  // tiles bounded by registers gain nothing from the split.  Strided FP64
  // tiles of long lines split too, except 2-line tiles of 8 points per thread
  // (64 KiB paired: two 512-thread workgroups per CU either way).
  constexpr bool SPLIT = (!F32 && LAYOUT == PACG_L_CONTIG && P >= 2048 && E > 8) || (LONG_S && (E > 8 || LINES > 2));
  // The tiles hold (re, im) pairs: one 8- or 16-byte LDS access per point and
  // per twiddle instead of two (the slot indices of the layouts above are
  // unchanged).  At 256^3 FP32 the fused pass goes from 79.5 to 76 us, with
  // 58 instead of 77 VGPRs (8 waves per SIMD instead of 6); FP64 tiles gain a
  // few percent at 2^16 and nothing at 256^3 or 2048^2.  The split tiles
  // above keep one array of one component.
  constexpr bool PAIRS = !SPLIT;
  using V2T = typename Vec2<T>::type;
  __shared__ T xr[NEED_X && !PAIRS ? LDS_N : 1];
  __shared__ T xi[NEED_X && !SPLIT && !PAIRS ? LDS_N : 1];
  __shared__ V2T xv[NEED_X && PAIRS ? LDS_N : 1];
  // w_P in LDS: the whole table up to 512 points, two levels above
  // (w_P^p = hi[p >> 6] * lo[p & 63]) so that long lines keep LDS for data.
  // The 64 lo entries sit after the P / 64 hi entries with one pad slot per
  // 8: the phases whose twiddle exponents step by 8 read lo entries 8 apart,
  // which would otherwise fall into one bank group (an 8-way conflict with
  // 16-byte entries).
  constexpr bool TW2 = P > 512;
  constexpr int TWN = TW2 ? P / 64 + 72 : P;
  constexpr int TWLO = P / 64;  // first lo slot
  auto lo_slot = [](int j) PACG_INL { return TWLO + j + (j >> 3); };
  __shared__ T twr[PAIRS ? 1 : TWN], twi[PAIRS ? 1 : TWN];
  __shared__ V2T twv[PAIRS ? TWN : 1];

  const int tid = threadIdx.x;
  int tl, ll;
  if constexpr (LAYOUT == PACG_L_STRIDED) {
    ll = tid % LINES;
    tl = tid / LINES;
  } else {
    ll = tid / TPL;
    tl = tid % TPL;
  }
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t S = 1u << a.log2_S;
  uint32_t line_base;  // element offset of point 0 of this thread's line
  uint32_t jpos = 0;   // j of this line
  bool valid;
  if constexpr (LAYOUT == PACG_L_STRIDED) {
    const uint32_t lg = tile * LINES + ll;
    const uint32_t o = lg >> a.log2_S;
    jpos = lg & (S - 1u);
    valid = o < a.n_o;
    line_base = o * (uint32_t)P * S + jpos;
  } else {
    const uint32_t line = tile * LINES + ll;
    valid = line < a.n_o;
    line_base = line * (uint32_t)P;
  }
  const uint32_t stride = (LAYOUT == PACG_L_STRIDED) ? S : 1u;
  // Lines past the end (in the last tile only) read line 0 and store nothing:
  // loads need no predicate, so no branch per element and no zero fill.
  const uint32_t ld_base = valid ? line_base : 0u;

  T vr[E], vi[E];

  // w_P table -> LDS
  {
    using V2 = typename Vec2<T>::type;
    const V2* tp = reinterpret_cast<const V2*>(a.tw_p);
    constexpr int NTW = TW2 ? P / 64 + 64 : P;  // entries (the pad slots stay unwritten)
    for (int i = tid; i < NTW; i += NT) {
      const V2 w = tp[TW2 ? (i < P / 64 ? i * 64 : i - P / 64) : i];
      const int q = TW2 && i >= P / 64 ? lo_slot(i - P / 64) : i;
      if constexpr (PAIRS) {
        twv[q] = w;
      } else {
        twr[q] = w.x;
        twi[q] = w.y;
      }
    }
  }
  auto ptw = [&](auto PH_, auto CJ_) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr bool CJ = decltype(CJ_)::value;
    // FP64 strided tiles of long lines with 16-point phases: one twiddle at a
    // time (with all 30 table reads of a phase hoisted they spill to scratch)
    constexpr int SER = (LONG_S && E >= 16) ? 2 : 0;
    if constexpr (PAIRS && TW2) {
      phase_twiddle_f<T, SH, PH, CJ, SER>(vr, vi, tl, [&](int idx, T& wr, T& wi) PACG_INL {
        const V2T h = twv[idx >> 6], l = twv[lo_slot(idx & 63)];
        wr = h.x * l.x - h.y * l.y;
        wi = h.x * l.y + h.y * l.x;
      });
    } else if constexpr (PAIRS) {
      phase_twiddle_f<T, SH, PH, CJ>(vr, vi, tl, [&](int idx, T& wr, T& wi) PACG_INL {
        const V2T w = twv[idx];
        wr = w.x;
        wi = w.y;
      });
    } else if constexpr (TW2) {
      phase_twiddle_f<T, SH, PH, CJ, SER>(vr, vi, tl, [&](int idx, T& wr, T& wi) PACG_INL {
        const T hr = twr[idx >> 6], hi = twi[idx >> 6];
        const T lr = twr[lo_slot(idx & 63)], li = twi[lo_slot(idx & 63)];
        wr = hr * lr - hi * li;
        wi = hr * li + hi * lr;
      });
    } else {
      phase_twiddle<T, SH, PH, CJ>(vr, vi, twr, twi, tl);
    }
  };

  const T* __restrict__ inr = reinterpret_cast<const T*>(a.in_re);
  const T* __restrict__ ini = reinterpret_cast<const T*>(a.in_im);
  T* __restrict__ outr = reinterpret_cast<T*>(a.out_re);
  T* __restrict__ outi = reinterpret_cast<T*>(a.out_im);

  // LDS slot of element (c, r) of this thread in phase PH.  Every layout is
  // linear over GF(2) in the point index, and pos(tl, c, r) = pos(tl, 0, 0) |
  // pos(0, c, r) with disjoint bits, so the slot is a per-thread part (the same
  // for all elements of a phase) combined with a compile-time constant.
  auto lds_slot = [&](auto PH_, auto C_, auto R_) PACG_INL -> int {
    constexpr int PH = decltype(PH_)::value;
    const int st = pos<SH, PH>(tl, 0, 0);
    const int sc = pos<SH, PH>(0, decltype(C_)::value, decltype(R_)::value);
    if constexpr (LMODE == LM_ROWSWZ) {
      constexpr int NB = LINES == 16 ? 1 : 2, LGL = LINES == 16 ? 4 : 3;
      const int pat = (sc ^ row_flip<LINES>(sc)) & ((1 << NB) - 1);
      const int rest = (sc >> NB) << (NB + LGL);
      return ((((st ^ row_flip<LINES>(st)) ^ pat) << LGL) | ll) + rest;
    } else if constexpr (LMODE == LM_LINESWZ) {
      return ((ll * P) | lds_swz(st)) ^ lds_swz(sc);
    } else if constexpr (LMODE == LM_QUADSWZ) {
      return (ll * P) | (quad_swz(st) ^ quad_swz(sc));
    } else if constexpr (LMODE == LM_ROWXOR) {
      // row_fold is linear and ll < LINES: the constant part is one XOR
      return (row_fold<XNB, SH::LGP>(st) * LINES + ll) ^ (row_fold<XNB, SH::LGP>(sc) * LINES);
    } else if constexpr (LAYOUT == PACG_L_STRIDED) {
      return (st | sc) * LDJ + ll;
    } else {
      const int sp = st | sc;
      return ll * LDP + sp + (sp >> 3);
    }
  };

  // Phase 0 of a contiguous line holds R consecutive points per group:
  // move them as 16-byte vectors (VW elements) when the base pointers allow.
  constexpr int VW = 16 / (int)sizeof(T);
  constexpr bool VEC0 = (LAYOUT == PACG_L_CONTIG) && ((1 << SH::lgR(0)) % VW == 0);
  using V4 = typename Vec4<T>::type;
  auto load_phase = [&](auto PH_, const T* __restrict__ br, const T* __restrict__ bi) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr int R = 1 << SH::lgR(PH);
    if constexpr (PH == 0 && VEC0) {
      if (a.vec16) {
        sfor<E / R>([&](auto c_) PACG_INL {
          constexpr int c = decltype(c_)::value;
          const uint32_t off = ld_base + (uint32_t)pos<SH, 0>(tl, c, 0);
          sfor<R / VW>([&](auto v_) PACG_INL {
            constexpr int v = decltype(v_)::value;
            V4 xr4, xi4;
            if constexpr ((NTM & NT_IN) != 0) {
              xr4 = Vec4<T>::load_nt(br + off + v * VW);
              xi4 = Vec4<T>::load_nt(bi + off + v * VW);
            } else {
              xr4 = reinterpret_cast<const V4*>(br + off)[v];
              xi4 = reinterpret_cast<const V4*>(bi + off)[v];
            }
            Vec4<T>::split(xr4, &vr[c * R + v * VW]);
            Vec4<T>::split(xi4, &vi[c * R + v * VW]);
          });
        });
        return;
      }
    }
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int s = pos<SH, PH>(tl, c, r);
        const uint32_t off = ld_base + (uint32_t)s * stride;
        if constexpr ((NTM & NT_IN) != 0) {
          vr[c * R + r] = __builtin_nontemporal_load(br + off);
          vi[c * R + r] = __builtin_nontemporal_load(bi + off);
        } else {
          vr[c * R + r] = br[off];
          vi[c * R + r] = bi[off];
        }
      });
    });
  };
  auto store_impl = [&](auto PH_, auto SCALED_, T scale) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr bool SCALED = decltype(SCALED_)::value;
    constexpr int R = 1 << SH::lgR(PH);
    // one branch around all the stores of the phase (not one per element)
    if (!valid) return;
    if constexpr (PH == 0 && VEC0) {
      if (a.vec16) {
        sfor<E / R>([&](auto c_) PACG_INL {
          constexpr int c = decltype(c_)::value;
          const uint32_t off = line_base + (uint32_t)pos<SH, 0>(tl, c, 0);
          sfor<R / VW>([&](auto v_) PACG_INL {
            constexpr int v = decltype(v_)::value;
            T tr[VW], ti[VW];
#pragma unroll
            for (int u = 0; u < VW; ++u) {
              tr[u] = SCALED ? vr[c * R + v * VW + u] * scale : vr[c * R + v * VW + u];
              ti[u] = SCALED ? vi[c * R + v * VW + u] * scale : vi[c * R + v * VW + u];
            }
            reinterpret_cast<V4*>(outr + off)[v] = Vec4<T>::join(tr);
            reinterpret_cast<V4*>(outi + off)[v] = Vec4<T>::join(ti);
          });
        });
        return;
      }
    }
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int s = pos<SH, PH>(tl, c, r);
        const uint32_t off = line_base + (uint32_t)s * stride;
        outr[off] = SCALED ? vr[c * R + r] * scale : vr[c * R + r];
        outi[off] = SCALED ? vi[c * R + r] * scale : vi[c * R + r];
      });
    });
  };
  // the multiply is skipped when the scale is 1 (every pass but the last of a
  // convolution or an inverse transform)
  auto store_phase = [&](auto PH_, T scale) PACG_INL {
    if (scale == T(1))
      store_impl(PH_, std::false_type{}, scale);
    else
      store_impl(PH_, std::true_type{}, scale);
  };
  // Contiguous lines of at most 64 threads lie inside one wavefront (tl is
  // tid % TPL, TPL a power of two), and every slot an exchange writes or reads
  // belongs to the thread's own line: the wave only has to wait for its own
  // LDS writes before reading (LDS operations of one wave complete in order),
  // not for the other waves of the workgroup.
  constexpr bool WAVE_X = LAYOUT == PACG_L_CONTIG && TPL <= 64;
  auto xsync = [&]() PACG_INL {
    if constexpr (WAVE_X)
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    else
      __syncthreads();
  };
  auto exchange = [&](auto FROM_, auto TO_) PACG_INL {
    constexpr int FROM = decltype(FROM_)::value;
    constexpr int TO = decltype(TO_)::value;
    constexpr int RF = 1 << SH::lgR(FROM);
    constexpr int RT = 1 << SH::lgR(TO);
    if constexpr (SPLIT) {
      auto move = [&](T* v) PACG_INL {
        xsync();
        sfor<E / RF>([&](auto c_) PACG_INL {
          constexpr int c = decltype(c_)::value;
          sfor<RF>([&](auto r_) PACG_INL {
            constexpr int r = decltype(r_)::value;
            xr[lds_slot(FROM_, c_, r_)] = v[c * RF + r];
          });
        });
        xsync();
        sfor<E / RT>([&](auto c_) PACG_INL {
          constexpr int c = decltype(c_)::value;
          sfor<RT>([&](auto r_) PACG_INL {
            constexpr int r = decltype(r_)::value;
            v[c * RT + r] = xr[lds_slot(TO_, c_, r_)];
          });
        });
      };
      move(vr);
      move(vi);
      return;
    }
    if constexpr (PAIRS) {
      xsync();
      sfor<E / RF>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<RF>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          V2T w;
          w.x = vr[c * RF + r];
          w.y = vi[c * RF + r];
          xv[lds_slot(FROM_, c_, r_)] = w;
        });
      });
      xsync();
      sfor<E / RT>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<RT>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const V2T w = xv[lds_slot(TO_, c_, r_)];
          vr[c * RT + r] = w.x;
          vi[c * RT + r] = w.y;
        });
      });
      return;
    }
    xsync();
    sfor<E / RF>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<RF>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int li = lds_slot(FROM_, c_, r_);
        xr[li] = vr[c * RF + r];
        xi[li] = vi[c * RF + r];
      });
    });
    xsync();
    sfor<E / RT>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<RT>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int li = lds_slot(TO_, c_, r_);
        vr[c * RT + r] = xr[li];
        vi[c * RT + r] = xi[li];
      });
    });
  };
  // pass-external twiddle w_K^{j' k(s)} on the phase-0 layout
  auto ext_twiddle = [&](bool cj) PACG_INL {
    using V2 = typename Vec2<T>::type;
    const V2* hi = reinterpret_cast<const V2*>(a.tw_hi);
    const V2* lo = reinterpret_cast<const V2*>(a.tw_lo);
    const uint32_t jp = jpos >> a.log2_inner;
    const uint32_t mask = (1u << a.lo_bits) - 1u;
    constexpr int R0 = 1 << SH::lgR(0);
    sfor<E / R0>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R0>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int s = pos<SH, 0>(tl, c, r);
        const uint32_t e = (jp * (uint32_t)rev_pass<SH>(s)) << a.ext_shift;
        const V2 h = hi[e >> a.lo_bits];
        const V2 l = lo[e & mask];
        T wr = h.x * l.x - h.y * l.y;
        T wi = h.x * l.y + h.y * l.x;
        if (cj) wi = -wi;
        cmul<T>(vr[c * R0 + r], vi[c * R0 + r], wr, wi);
        // FP64 16-point phases: one element at a time (its two table loads
        // per element, all hoisted, spill)
        if constexpr (LONG_S && E >= 16) __builtin_amdgcn_sched_barrier(0);
      });
    });
  };

  // fused pass: the spectrum at the phase-0 positions, loaded after the last
  // forward exchange, so its latency hides behind the phase-0 butterflies.
  // Loaded up front it would stay live through every forward phase: 81 instead
  // of 76 VGPRs for the 256-point FP32 tile, one wave per SIMD less, 3 % slower.
  constexpr int R0 = 1 << SH::lgR(0);
  TS hr_[KIND == PACG_K_FUSED ? E : 1], hi_[KIND == PACG_K_FUSED ? E : 1];
  auto load_spec = [&]() PACG_INL {
    if constexpr (KIND == PACG_K_FUSED) {
      const TS* __restrict__ hr = reinterpret_cast<const TS*>(a.spec_re);
      const TS* __restrict__ hi = reinterpret_cast<const TS*>(a.spec_im);
      sfor<E / R0>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<R0>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const int s = pos<SH, 0>(tl, c, r);
          const uint32_t off = ld_base + (uint32_t)s * stride;
          if constexpr ((NTM & NT_SPEC) != 0) {
            hr_[c * R0 + r] = __builtin_nontemporal_load(hr + off);
            hi_[c * R0 + r] = __builtin_nontemporal_load(hi + off);
          } else {
            hr_[c * R0 + r] = hr[off];
            hi_[c * R0 + r] = hi[off];
          }
        });
      });
    }
  };

  if constexpr (KIND == PACG_K_DIF || KIND == PACG_K_FUSED) {
    load_phase(std::integral_constant<int, NP - 1>{}, inr, ini);
    __syncthreads();  // twiddle table visible
    sfor<NP>([&](auto i_) PACG_INL {
      constexpr int PH = NP - 1 - decltype(i_)::value;
      if constexpr (PH < NP - 1) exchange(std::integral_constant<int, PH + 1>{}, std::integral_constant<int, PH>{});
      if constexpr (PH == 0) load_spec();
      phase_stages<T, SH, PH, M_DIF>(vr, vi);
      ptw(std::integral_constant<int, PH>{}, std::false_type{});
    });
    if constexpr (KIND == PACG_K_DIF) {
      if (a.ext) ext_twiddle(false);
      store_phase(std::integral_constant<int, 0>{}, (T)a.out_scale);
    } else {
      // spectral product at the same positions: x <- x * h; the factor
      // a.scale (1/n, a power of two for one convolution) is applied with
      // the output scale at the store, where it costs no extra multiply
#pragma unroll
      for (int e = 0; e < E; ++e) cmul<T>(vr[e], vi[e], (T)hr_[e], (T)hi_[e]);
      sfor<NP>([&](auto i_) PACG_INL {
        constexpr int PH = decltype(i_)::value;
        if constexpr (PH > 0) exchange(std::integral_constant<int, PH - 1>{}, std::integral_constant<int, PH>{});
        ptw(std::integral_constant<int, PH>{}, std::true_type{});
        phase_stages<T, SH, PH, M_DITC>(vr, vi);
      });
      store_phase(std::integral_constant<int, NP - 1>{}, (T)(a.scale * a.out_scale));
    }
  } else {
    constexpr int MODE = (KIND == PACG_K_DITC) ? M_DITC : M_DIT;
    load_phase(std::integral_constant<int, 0>{}, inr, ini);
    __syncthreads();
    if constexpr ((NTM & NO_EXT) == 0) {
      if (a.ext) ext_twiddle(MODE == M_DITC);
    }
    sfor<NP>([&](auto i_) PACG_INL {
      constexpr int PH = decltype(i_)::value;
      if constexpr (PH > 0) exchange(std::integral_constant<int, PH - 1>{}, std::integral_constant<int, PH>{});
      ptw(std::integral_constant<int, PH>{}, std::integral_constant<bool, MODE == M_DITC>{});
      phase_stages<T, SH, PH, MODE>(vr, vi);
    });
    store_phase(std::integral_constant<int, NP - 1>{}, (T)a.out_scale);
  }
}

template <typename T, typename TS, class SH, int KIND, int LAYOUT, int LINES>
int launch_pass(const pacg_pass_args_t* a, void* stream) {
  uint32_t blocks;
  if (LAYOUT == PACG_L_STRIDED) {
    const uint64_t lines = (uint64_t)a->n_o << a->log2_S;
    blocks = (uint32_t)((lines + LINES - 1) / LINES);
  } else {
    blocks = (a->n_o + LINES - 1) / LINES;
  }
  if (blocks == 0) return 0;
  // FP64-stored spectra stream past the caches when the plan asks for it
  // (strided tiles of an array and a spectrum that exceed the Infinity Cache
  // together); so does the input of a convolution's first A^T pass when the
  // plan's autotuner found that faster
  if constexpr (KIND == PACG_K_FUSED && sizeof(TS) == 8 && LAYOUT == PACG_L_STRIDED) {
    if (a->spec_nt) {
      hipLaunchKernelGGL((pass_kernel<T, TS, SH, KIND, LAYOUT, LINES, NT_SPEC>), dim3(blocks), dim3(LINES * SH::TPL), 0,
                         (hipStream_t)stream, *a);
      return (int)hipGetLastError();
    }
  }
  if constexpr ((KIND == PACG_K_DITC || KIND == PACG_K_DIT) && sizeof(T) == 8 && LAYOUT == PACG_L_STRIDED &&
                SH::P >= 2048) {
    if (!a->ext) {
      hipLaunchKernelGGL((pass_kernel<T, TS, SH, KIND, LAYOUT, LINES, NO_EXT>), dim3(blocks), dim3(LINES * SH::TPL), 0,
                         (hipStream_t)stream, *a);
      return (int)hipGetLastError();
    }
  }
  if constexpr (KIND == PACG_K_DIF) {
    if (a->in_nt) {
      hipLaunchKernelGGL((pass_kernel<T, TS, SH, KIND, LAYOUT, LINES, NT_IN>), dim3(blocks), dim3(LINES * SH::TPL), 0,
                         (hipStream_t)stream, *a);
      return (int)hipGetLastError();
    }
  }
  hipLaunchKernelGGL((pass_kernel<T, TS, SH, KIND, LAYOUT, LINES>), dim3(blocks), dim3(LINES * SH::TPL), 0,
                     (hipStream_t)stream, *a);
  return (int)hipGetLastError();
}

}  // namespace pacg
