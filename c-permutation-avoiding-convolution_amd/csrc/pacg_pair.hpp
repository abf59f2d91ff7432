// Pair kernels: passes of two different axes in one HBM round trip.
//
// A tile holds PX x PY x U points: PX points of pass X (axis a, SX elements
// apart), PY points of pass Y (axis b > a, SY apart) and U neighbouring
// columns at unit stride.  Pass X runs on the PY*U X-lines of the tile, pass
// Y on its PX*U Y-lines, and the two hand their points over through LDS.
// Stages of different axes commute, so a pair is exactly the two passes run
// one after the other (reference butterflies_3d.c applies the axes one at a
// time; the product is the same):
//   A^T (DIF):      X stages, X external twiddle, Y stages
//   A, conj(A):     Y stages, X external twiddle, X stages
//   fused:          X DIF, Y DIF, * spectrum * scale, Y conj(A), X conj(A)
// The fused form is used only when both passes are the bottom passes of their
// axes (no external twiddle), where it is the last A^T step, the spectral
// product (convolution_3d.c:58-64) and the first conj(A) step.
//
// Layout choices:
//   * YC: Y-lines are contiguous rows (SY == 1, U == 1), moved as 16-byte
//     vectors; otherwise a thread may hold G = E / PY whole Y-lines (a single
//     register phase), which lets a short bottom pass sit beside U = 16
//     columns (64-byte runs);
//   * the LDS tile is [PX][PY*U], XOR-swizzled on the column index so that
//     the Y-lines of one wave fall on distinct banks (padded rows for YC);
//   * the next tile's input is loaded while the current tile is transformed
//     (PF), except where those registers would spill;
//   * workgroups are persistent, walk a contiguous range of tiles per XCD and
//     load the next tile's input into registers while the current one is in
//     the butterflies;
//   * the external twiddles of a tile (one X-line offset per tile) are formed
//     once per tile in LDS; HBM addresses are 32-bit offsets from a
//     workgroup-uniform tile base.
#pragma once

#include "pacg_pass.hpp"

namespace pacg {

template <class SX, class SY, int U, bool YC>
struct PairGeom {
  static constexpr int E = SX::E, PX = SX::P, PY = SY::P;
  static constexpr int G = E / SY::E;  // Y-lines per thread
  static constexpr int NLX = PY * U;
  static constexpr int NLY = PX * U;
  static constexpr int NT = NLX * SX::TPL;
  static constexpr int RW = PY * U;
  static constexpr int ROW = YC ? (PY + PY / 32 + 1) : RW;
  static constexpr int LDS_N = PX * ROW;
  // Column tiles: a wave's Y-lines are U columns of 64 / U consecutive rows x;
  // XOR-ing the row's column index with a multiple of U spreads those rows
  // over distinct banks (64 four-byte banks).
  static constexpr int XPW = U >= 64 ? 1 : 64 / U;
  static constexpr int RPB = RW >= 64 ? 1 : 64 / RW;  // rows per 64 banks
  static constexpr int SWM = (XPW / RPB) < (RW / U) ? (XPW / RPB) : (RW / U);
  // LDS slot of tile point (x, y, u)
  static __device__ __forceinline__ int lds(int x, int y, int u) {
    if constexpr (YC) {
      return x * ROW + y + (y >> 5);
    } else if constexpr (SWM > 1) {
      return x * ROW + ((y * U + u) ^ (((x / RPB) & (SWM - 1)) * U));
    } else {
      return x * ROW + y * U + u;
    }
  }
};

// PF: load the next tile's input while the current one is transformed
// (off where the extra registers would spill)
template <typename T, typename TS, class SX, class SY, int U, bool YC, int KIND, bool PF>
__global__ void __launch_bounds__((PairGeom<SX, SY, U, YC>::NT)) pair_kernel(const pacg_pair_args_t a) {
  using PG = PairGeom<SX, SY, U, YC>;
  static_assert(SX::valid() && SY::valid(), "invalid pass shape");
  static_assert(!YC || U == 1, "contiguous Y rows have no extra columns");
  constexpr int E = PG::E, PX = PG::PX, PY = PG::PY, G = PG::G, EY = SY::E;
  constexpr int NLX = PG::NLX, NT = PG::NT;
  static_assert(G >= 1 && G * EY == E, "points per thread");
  static_assert(G == 1 || (!YC && SY::NP == 1 && SY::TPL == 1), "several Y-lines per thread need one-phase Y");
  static_assert(NT * G == PG::NLY * SY::TPL && NT <= 1024, "thread count");
  constexpr int NPX = SX::NP, NPY = SY::NP;
  constexpr int XT = PX < 64 ? 64 : PX;  // external twiddle slots (one per X position)
  __shared__ T xr[PG::LDS_N];
  __shared__ T xi[PG::LDS_N];
  __shared__ T twxr[PX], twxi[PX], twyr[PY], twyi[PY], ter[XT], tei[XT];

  // The thread index is re-read opaquely at the top of every tile (see the
  // loop) so that the per-thread addresses below are formed per tile rather
  // than hoisted out of the loop, where they would take registers for good.
  int tid = threadIdx.x;
  const int tid0 = tid;
#define llx (tid % NLX)
#define tlx (tid / NLX)
  // Y-line of group g and the thread's index within it
  auto lly_of = [&](int g) PACG_INL -> int {
    if constexpr (G > 1)
      return tid + g * NT;
    else if constexpr (YC)
      return tid / SY::TPL;
    else
      return tid % PG::NLY;
  };
#define tly ((G > 1) ? 0 : (YC ? tid % SY::TPL : tid / PG::NLY))

  const uint32_t SXs = 1u << a.log2_SX, SYs = 1u << a.log2_SY;
  const uint32_t ystride = YC ? 1u : SYs;
  // offsets (elements) of this thread's X-line and Y-lines from the tile base
#define xoff ((uint32_t)(llx / U) * SYs + (uint32_t)(llx % U))
  auto yoff_of = [&](int g) PACG_INL -> uint32_t {
    const int l = lly_of(g);
    if constexpr (YC)
      return (uint32_t)l * SXs;
    else
      return (uint32_t)(l / U) * SXs + (uint32_t)(l % U);
  };
  auto lds_x = [&](int s) PACG_INL -> int { return PG::lds(s, llx / U, llx % U); };
  auto lds_y = [&](int g, int s) PACG_INL -> int {
    const int l = lly_of(g);
    if constexpr (YC)
      return PG::lds(l, s, 0);
    else
      return PG::lds(l / U, s, l % U);
  };

  // ---- twiddle tables, once per workgroup ----
  {
    using V2 = typename Vec2<T>::type;
    const V2* tx = reinterpret_cast<const V2*>(a.tw_x);
    for (int i = tid0; i < PX; i += NT) {
      V2 w = tx[i];
      twxr[i] = w.x;
      twxi[i] = w.y;
    }
    const V2* ty = reinterpret_cast<const V2*>(a.tw_y);
    for (int i = tid0; i < PY; i += NT) {
      V2 w = ty[i];
      twyr[i] = w.x;
      twyi[i] = w.y;
    }
  }

  // ---- tiles of this workgroup: a contiguous range per XCD when the counts allow ----
  uint32_t t0, tend, tstep;
  {
    const uint32_t nt = a.n_tiles, nb = gridDim.x, b = blockIdx.x;
    if ((nb & 7u) == 0u && (nt & 7u) == 0u && nb <= nt) {
      const uint32_t per = nt >> 3, k = b & 7u, j = b >> 3;
      t0 = k * per + j;
      tend = (k + 1u) * per;
      tstep = nb >> 3;
    } else {
      t0 = b;
      tend = nt;
      tstep = nb;
    }
  }
  auto tile_base = [&](uint32_t tile) PACG_INL -> uint32_t {
    const uint32_t ub = tile & ((1u << a.log2_nub) - 1u);
    const uint32_t rest = tile >> a.log2_nub;
    const uint32_t oy = rest & ((1u << a.log2_noy) - 1u);
    const uint32_t ox = rest >> a.log2_noy;
    return ox * (uint32_t)PX * SXs + oy * (uint32_t)PY * SYs + ub * (uint32_t)U;
  };
  // element at a 32-bit offset from a tile base pointer
  auto at = [](const T* p, uint32_t off) PACG_INL -> const T& {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (off * (uint32_t)sizeof(T)));
  };
  auto atw = [](T* p, uint32_t off) PACG_INL -> T& {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + (off * (uint32_t)sizeof(T)));
  };

  T vr[E], vi[E];
  T pr[PF ? E : 1], pi[PF ? E : 1];  // next tile's input

  constexpr int VW = 16 / (int)sizeof(T);
  constexpr bool VECY = YC && ((1 << SY::lgR(0)) % VW == 0);
  using V4 = typename Vec4<T>::type;

  // ---- HBM -> registers ----
  auto load_x = [&](auto PH_, const T* br, const T* bi, T* dr, T* di) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr int R = 1 << SX::lgR(PH);
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const uint32_t off = xoff + (uint32_t)pos<SX, PH>(tlx, c, r) * SXs;
        dr[c * R + r] = at(br, off);
        di[c * R + r] = at(bi, off);
      });
    });
  };
  auto load_y0 = [&](const T* br, const T* bi, T* dr, T* di) PACG_INL {
    constexpr int R = 1 << SY::lgR(0);
    if constexpr (VECY) {
      if (a.vec16) {
        const uint32_t yo = yoff_of(0);
        sfor<E / R>([&](auto c_) PACG_INL {
          constexpr int c = decltype(c_)::value;
          const uint32_t off = yo + (uint32_t)pos<SY, 0>(tly, c, 0);
          sfor<R / VW>([&](auto v_) PACG_INL {
            constexpr int v = decltype(v_)::value;
            const V4 xr4 = *reinterpret_cast<const V4*>(&at(br, off + v * VW));
            const V4 xi4 = *reinterpret_cast<const V4*>(&at(bi, off + v * VW));
            Vec4<T>::split(xr4, &dr[c * R + v * VW]);
            Vec4<T>::split(xi4, &di[c * R + v * VW]);
          });
        });
        return;
      }
    }
    sfor<G>([&](auto g_) PACG_INL {
      constexpr int g = decltype(g_)::value;
      const uint32_t yo = yoff_of(g);
      sfor<EY / R>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<R>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const uint32_t off = yo + (uint32_t)pos<SY, 0>(tly, c, r) * ystride;
          dr[g * EY + c * R + r] = at(br, off);
          di[g * EY + c * R + r] = at(bi, off);
        });
      });
    });
  };
  auto load_in = [&](uint32_t tile, T* dr, T* di) PACG_INL {
    const uint32_t base = tile_base(tile);
    const T* br = reinterpret_cast<const T*>(a.in_re) + base;
    const T* bi = reinterpret_cast<const T*>(a.in_im) + base;
    if constexpr (KIND == PACG_K_DIF || KIND == PACG_K_FUSED)
      load_x(std::integral_constant<int, NPX - 1>{}, br, bi, dr, di);
    else
      load_y0(br, bi, dr, di);
  };

  // ---- registers -> HBM ----
  auto store_x = [&](T* br, T* bi, T scale) PACG_INL {
    constexpr int PH = NPX - 1;
    constexpr int R = 1 << SX::lgR(PH);
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const uint32_t off = xoff + (uint32_t)pos<SX, PH>(tlx, c, r) * SXs;
        atw(br, off) = vr[c * R + r] * scale;
        atw(bi, off) = vi[c * R + r] * scale;
      });
    });
  };
  auto store_y0 = [&](T* br, T* bi, T scale) PACG_INL {
    constexpr int R = 1 << SY::lgR(0);
    if constexpr (VECY) {
      if (a.vec16) {
        const uint32_t yo = yoff_of(0);
        sfor<E / R>([&](auto c_) PACG_INL {
          constexpr int c = decltype(c_)::value;
          const uint32_t off = yo + (uint32_t)pos<SY, 0>(tly, c, 0);
          sfor<R / VW>([&](auto v_) PACG_INL {
            constexpr int v = decltype(v_)::value;
            T tr[VW], ti[VW];
#pragma unroll
            for (int w = 0; w < VW; ++w) {
              tr[w] = vr[c * R + v * VW + w] * scale;
              ti[w] = vi[c * R + v * VW + w] * scale;
            }
            *reinterpret_cast<V4*>(&atw(br, off + v * VW)) = Vec4<T>::join(tr);
            *reinterpret_cast<V4*>(&atw(bi, off + v * VW)) = Vec4<T>::join(ti);
          });
        });
        return;
      }
    }
    sfor<G>([&](auto g_) PACG_INL {
      constexpr int g = decltype(g_)::value;
      const uint32_t yo = yoff_of(g);
      sfor<EY / R>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<R>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const uint32_t off = yo + (uint32_t)pos<SY, 0>(tly, c, r) * ystride;
          atw(br, off) = vr[g * EY + c * R + r] * scale;
          atw(bi, off) = vi[g * EY + c * R + r] * scale;
        });
      });
    });
  };

  // ---- registers <-> LDS ----
  auto put_x = [&](auto PH_) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr int R = 1 << SX::lgR(PH);
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int li = lds_x(pos<SX, PH>(tlx, c, r));
        xr[li] = vr[c * R + r];
        xi[li] = vi[c * R + r];
      });
    });
  };
  auto get_x = [&](auto PH_) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr int R = 1 << SX::lgR(PH);
    sfor<E / R>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int li = lds_x(pos<SX, PH>(tlx, c, r));
        vr[c * R + r] = xr[li];
        vi[c * R + r] = xi[li];
      });
    });
  };
  auto put_y = [&](auto PH_) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr int R = 1 << SY::lgR(PH);
    sfor<G>([&](auto g_) PACG_INL {
      constexpr int g = decltype(g_)::value;
      sfor<EY / R>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<R>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const int li = lds_y(g, pos<SY, PH>(tly, c, r));
          xr[li] = vr[g * EY + c * R + r];
          xi[li] = vi[g * EY + c * R + r];
        });
      });
    });
  };
  auto get_y = [&](auto PH_) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr int R = 1 << SY::lgR(PH);
    sfor<G>([&](auto g_) PACG_INL {
      constexpr int g = decltype(g_)::value;
      sfor<EY / R>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<R>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const int li = lds_y(g, pos<SY, PH>(tly, c, r));
          vr[g * EY + c * R + r] = xr[li];
          vi[g * EY + c * R + r] = xi[li];
        });
      });
    });
  };
  using I0 = std::integral_constant<int, 0>;
  auto x_to_y = [&]() PACG_INL {
    __syncthreads();
    put_x(I0{});
    __syncthreads();
    get_y(std::integral_constant<int, NPY - 1>{});
  };
  auto y_to_x = [&]() PACG_INL {
    __syncthreads();
    put_y(std::integral_constant<int, NPY - 1>{});
    __syncthreads();
    get_x(I0{});
  };

  // ---- the passes ----
  auto x_dif = [&]() PACG_INL {
    sfor<NPX>([&](auto i_) PACG_INL {
      constexpr int PH = NPX - 1 - decltype(i_)::value;
      if constexpr (PH < NPX - 1) {
        __syncthreads();
        put_x(std::integral_constant<int, PH + 1>{});
        __syncthreads();
        get_x(std::integral_constant<int, PH>{});
      }
      phase_stages<T, SX, PH, M_DIF>(vr, vi);
      phase_twiddle<T, SX, PH, false>(vr, vi, twxr, twxi, tlx);
    });
  };
  auto y_stages = [&](auto PH_, auto MODE_) PACG_INL {
    constexpr int PH = decltype(PH_)::value;
    constexpr int MODE = decltype(MODE_)::value;
    sfor<G>([&](auto g_) PACG_INL {
      constexpr int g = decltype(g_)::value;
      phase_stages<T, SY, PH, MODE>(vr + g * EY, vi + g * EY);
    });
  };
  auto y_dif = [&]() PACG_INL {
    sfor<NPY>([&](auto i_) PACG_INL {
      constexpr int PH = NPY - 1 - decltype(i_)::value;
      if constexpr (PH < NPY - 1) {
        __syncthreads();
        put_y(std::integral_constant<int, PH + 1>{});
        __syncthreads();
        get_y(std::integral_constant<int, PH>{});
      }
      y_stages(std::integral_constant<int, PH>{}, std::integral_constant<int, M_DIF>{});
      if constexpr (G == 1) phase_twiddle<T, SY, PH, false>(vr, vi, twyr, twyi, tly);
    });
  };
  auto x_dit = [&](auto MODE_) PACG_INL {
    constexpr int MODE = decltype(MODE_)::value;
    sfor<NPX>([&](auto i_) PACG_INL {
      constexpr int PH = decltype(i_)::value;
      if constexpr (PH > 0) {
        __syncthreads();
        put_x(std::integral_constant<int, PH - 1>{});
        __syncthreads();
        get_x(std::integral_constant<int, PH>{});
      }
      phase_twiddle<T, SX, PH, MODE == M_DITC>(vr, vi, twxr, twxi, tlx);
      phase_stages<T, SX, PH, MODE>(vr, vi);
    });
  };
  auto y_dit = [&](auto MODE_) PACG_INL {
    constexpr int MODE = decltype(MODE_)::value;
    sfor<NPY>([&](auto i_) PACG_INL {
      constexpr int PH = decltype(i_)::value;
      if constexpr (PH > 0) {
        __syncthreads();
        put_y(std::integral_constant<int, PH - 1>{});
        __syncthreads();
        get_y(std::integral_constant<int, PH>{});
      }
      if constexpr (G == 1) phase_twiddle<T, SY, PH, MODE == M_DITC>(vr, vi, twyr, twyi, tly);
      y_stages(std::integral_constant<int, PH>{}, std::integral_constant<int, MODE>{});
    });
  };
  // external twiddles of pass X for one tile: w_K^{j' k(s)}, j' = the tile's
  // within-axis lower index (the same for every X-line of the tile: the Y
  // pass spans the axes inside X's axis).  Thread s < PX fetches entry s.
  static_assert(PX <= NT, "one thread per external twiddle");
  using V2 = typename Vec2<T>::type;
  auto ext_fetch = [&](uint32_t base, V2& h, V2& l) PACG_INL {
    const V2* hi = reinterpret_cast<const V2*>(a.tw_hi);
    const V2* lo = reinterpret_cast<const V2*>(a.tw_lo);
    const uint32_t jp = (base & (SXs - 1u)) >> a.log2_inner_x;
    const uint32_t e = (jp * (uint32_t)rev_pass<SX>(tid)) << a.ext_shift_x;
    h = hi[e >> a.lo_bits];
    l = lo[e & ((1u << a.lo_bits) - 1u)];
  };
  auto ext_apply = [&](bool cj) PACG_INL {
    constexpr int R0 = 1 << SX::lgR(0);
    sfor<E / R0>([&](auto c_) PACG_INL {
      constexpr int c = decltype(c_)::value;
      sfor<R0>([&](auto r_) PACG_INL {
        constexpr int r = decltype(r_)::value;
        const int s = pos<SX, 0>(tlx, c, r);
        T wi = tei[s];
        if (cj) wi = -wi;
        cmul<T>(vr[c * R0 + r], vi[c * R0 + r], ter[s], wi);
      });
    });
  };

  constexpr int RY0 = 1 << SY::lgR(0);
  TS hr_[KIND == PACG_K_FUSED ? E : 1], hi_[KIND == PACG_K_FUSED ? E : 1];
  // the spectrum at the Y phase-0 positions of one tile; fetched after the
  // forward X phases so that it is live only across the Y stages
  auto load_spectrum = [&](uint32_t base) PACG_INL {
    const TS* hr = reinterpret_cast<const TS*>(a.spec_re) + base;
    const TS* hi = reinterpret_cast<const TS*>(a.spec_im) + base;
    sfor<G>([&](auto g_) PACG_INL {
      constexpr int g = decltype(g_)::value;
      const uint32_t yo = yoff_of(g);
      sfor<EY / RY0>([&](auto c_) PACG_INL {
        constexpr int c = decltype(c_)::value;
        sfor<RY0>([&](auto r_) PACG_INL {
          constexpr int r = decltype(r_)::value;
          const uint32_t off = yo + (uint32_t)pos<SY, 0>(tly, c, r) * ystride;
          hr_[g * EY + c * RY0 + r] =
              *reinterpret_cast<const TS*>(reinterpret_cast<const char*>(hr) + off * (uint32_t)sizeof(TS));
          hi_[g * EY + c * RY0 + r] =
              *reinterpret_cast<const TS*>(reinterpret_cast<const char*>(hi) + off * (uint32_t)sizeof(TS));
        });
      });
    });
  };

  if (t0 < tend) load_in(t0, vr, vi);
  for (uint32_t tile = t0; tile < tend; tile += tstep) {
    tid = tid0;
    asm volatile("" : "+v"(tid));
    const uint32_t base = tile_base(tile);
    V2 eh{}, el{};
    if (a.ext_x && tid < PX) ext_fetch(base, eh, el);
    // the next tile's input: issued after this tile's own loads so that
    // waiting for those never waits for it
    const uint32_t tn = tile + tstep;
    if constexpr (PF) {
      if (tn < tend) load_in(tn, pr, pi);
    }
    __syncthreads();  // twiddle tables visible; the previous tile is done with LDS
    if (a.ext_x) {
      if (tid < PX) {
        ter[tid] = eh.x * el.x - eh.y * el.y;
        tei[tid] = eh.x * el.y + eh.y * el.x;
      }
      __syncthreads();
    }

    T* outr = reinterpret_cast<T*>(a.out_re) + base;
    T* outi = reinterpret_cast<T*>(a.out_im) + base;
    if constexpr (KIND == PACG_K_DIF || KIND == PACG_K_FUSED) {
      x_dif();
      if (a.ext_x) ext_apply(false);
      if constexpr (KIND == PACG_K_FUSED) load_spectrum(base);
      x_to_y();
      y_dif();
      if constexpr (KIND == PACG_K_DIF) {
        store_y0(outr, outi, (T)a.out_scale);
      } else {
        const T sc = (T)a.scale;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          cmul<T>(vr[e], vi[e], (T)hr_[e], (T)hi_[e]);
          vr[e] *= sc;
          vi[e] *= sc;
        }
        y_dit(std::integral_constant<int, M_DITC>{});
        y_to_x();
        if (a.ext_x) ext_apply(true);
        x_dit(std::integral_constant<int, M_DITC>{});
        store_x(outr, outi, (T)a.out_scale);
      }
    } else {
      constexpr int MODE = (KIND == PACG_K_DITC) ? M_DITC : M_DIT;
      y_dit(std::integral_constant<int, MODE>{});
      y_to_x();
      if (a.ext_x) ext_apply(MODE == M_DITC);
      x_dit(std::integral_constant<int, MODE>{});
      store_x(outr, outi, (T)a.out_scale);
    }
    if (tn < tend) {
      if constexpr (PF) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          vr[e] = pr[e];
          vi[e] = pi[e];
        }
      } else {
        tid = tid0;
        asm volatile("" : "+v"(tid));
        load_in(tn, vr, vi);
      }
    }
  }
#undef llx
#undef tlx
#undef tly
#undef xoff
}

// Resident workgroups per CU for one pair kernel (occupancy query, cached).
template <typename T, typename TS, class SX, class SY, int U, bool YC, int KIND, bool PF>
int pair_blocks_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pair_kernel<T, TS, SX, SY, U, YC, KIND, PF>,
                                                     PairGeom<SX, SY, U, YC>::NT, 0) != hipSuccess ||
        n < 1)
      n = 1;
    cached = n;
  }
  return cached;
}

template <typename T, typename TS, class SX, class SY, int U, bool YC, int KIND, bool PF>
int launch_pair(const pacg_pair_args_t* a, void* stream) {
  if (a->n_tiles == 0) return 0;
  uint32_t grid = (uint32_t)a->num_cus * (uint32_t)pair_blocks_per_cu<T, TS, SX, SY, U, YC, KIND, PF>();
  if (grid > a->n_tiles) grid = a->n_tiles;
  if (grid > 8u) grid &= ~7u;  // whole XCD groups
  hipLaunchKernelGGL((pair_kernel<T, TS, SX, SY, U, YC, KIND, PF>), dim3(grid), dim3(PairGeom<SX, SY, U, YC>::NT), 0,
                     (hipStream_t)stream, *a);
  return (int)hipGetLastError();
}

}  // namespace pacg
