// eikonal_kernels.hpp -- the model-templated kernels of the eikonal solver
// (tile update in its variants, precomputed-scheme export, flow export) and
// the per-model kernel tables.  Each model's instantiations live in their
// own translation unit (eik_<model>.hip) so that they compile in parallel;
// eikonal.hip holds the host side and the model-independent kernels.
//
// Tile update (reference: Update.h:20-227, HFMIter.h:10-86,
// Propagation.h:18-167, FiniteDifferences.h:17-106):
//   one workgroup per active tile, one lane per cell.  The lane loads its
//   value, rhs, seed bit and geometry (coalesced: the tiled layout makes a
//   tile contiguous), builds its stencil in registers (Selling for
//   Riemann/Rander) or reads it precomputed, then resolves each of its
//   2*nsym neighbours once: an LDS slot when the neighbour lies in the tile,
//   otherwise the neighbour's value gathered from HBM/L2 and held in a
//   register for the whole visit (values outside the tile do not change
//   during the in-tile sweeps).  The tile's own values live in LDS and are
//   relaxed niter_i times (Jacobi inside a sweep, early exit once a sweep
//   changes nothing -- which cannot alter the result since the next sweep
//   would read identical inputs).  The change test
//   |u_old-u_new| <= atol + rtol|u_new| is reduced with a block-wide OR (a
//   single wave64 ballot for 64-cell tiles) and drives tile activation.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.hpp"
#include "device.hpp"
#include "geodesic_kernels.hpp"
#include "worklist.hpp"

namespace agd {

// Source factorization parameters (see factor_shifts below).
template <class S> struct FactorDesc {
  S origin[3];
  S r2;     // squared radius; 0 disables
  S m[14];  // M: diagonal (Isotropic, Diagonal) or lower triangle; TTI:
            // [l0, l1, q00, q10, q11, A (row-major)]
  // TTI: the shifts of every cell of the ball, computed once per solve
  // (eik_tti_factor) for the box [lo, lo + n) of cells: component
  // (m * 2 NS + 2k + s) (then, with order 2, the further neighbours after
  // NMIX * 2 NS components) of box cell b at comp * nbox + b.
  const S* tab;
  int32_t lo[3], n[3];
  int64_t nbox;
};

// Index of the grid cell xt in the factor box of f, -1 outside the ball
// (the predicate of factor_shifts, FiniteDifferences.h:30-33).
template <class S, int D>
__device__ inline int64_t factor_box_index(const FactorDesc<S>& f, const int* xt) {
#pragma clang fp contract(off)
  S n2 = S(0);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const S x = S(xt[c]) - f.origin[c];
    n2 = fma_(x, x, n2);
  }
  if (!(n2 < f.r2)) return -1;
  int64_t b = 0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const int r = xt[c] - f.lo[c];
    if (r < 0 || r >= f.n[c]) return -1;
    b = b * f.n[c] + r;
  }
  return b;
}

// One instance of a batched solve: its own grid and constants; its tiles
// occupy [tile_base, tile_base + g.ntiles) of the batch's tile space.
struct InstDesc {
  GridDesc g;
  int32_t tile_base;
  int32_t pad;
  double iso_w[3];
  double atol, rtol;
};

template <class S> struct EikArgs {
  S* u;
  const S* geom;
  const S* sw;          // precomputed stencil weights (NS components) or null
  const int32_t* scode; // precomputed neighbour codes (2*NS components) or null
  const int32_t* scode2; // TTI, second order: further-neighbour codes, or null
  const S* rhs;
  const uint32_t* seeds;
  const uint8_t* wall;
  GridDesc g;
  S atol, rtol;
  S iso_w[3];
  int32_t niter_i;
  int32_t solver;
  int32_t fim_front;
  // worklist; FIM scores are packed (pass << 8) | score, two arrays by parity
  WorkCtx w;
  // batched solves: instance table and tile -> instance map
  const InstDesc* inst;
  const int32_t* tile_inst;
  S order2_thr;  // VAR_ORDER2: threshold of the second-order criterion
  FactorDesc<S> fac;  // VAR_GENERAL, VAR_ORDER2: source factorization
  BabCtx bab;         // bound_active_blocks (AGSI), ctl null: off
};

// HFM.h:91-134: where |u0 - 2v + v2| < thr * max_k(u0 - v_k)_+ and u0 > v,
// replace v by v + (v - v2)/3 and its weight w by 9/4 w.  Returns the mask
// of corrected neighbours (their flow weights are scaled by 3/2).
template <class S, int NS>
__device__ inline unsigned order2_correct(S u0, S thr, S (&v)[NS], const S (&v2)[NS], S (&w)[NS]) {
  S bound = S(0);
#pragma unroll
  for (int k = 0; k < NS; ++k) bound = max(bound, u0 - v[k]);
  unsigned mask = 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const S diff2 = abs_(u0 - S(2) * v[k] + v2[k]);
    if (diff2 < thr * bound && u0 > v[k]) {
      v[k] += (v[k] - v2[k]) / S(3);
      w[k] *= S(2.25);
      mask |= 1u << k;
    }
  }
  return mask;
}

// ------------------------------------------------------------------ factor
// Source factorization (EuclideanFactor.h:8-30, Constants.h:50-61,
// FiniteDifferences.h:30-45).  Inside the ball |x - origin|^2 < r2 (grid
// units) the value of the neighbour x + s e is shifted by
//   fact_s = T(x) - T(x + s e) + s <e, grad T(x)>,   T(y) = |y - origin|_M,
// so that T solves the scheme exactly when the metric is the constant M.
// The second-order correction reads its further neighbour x + 2 s e shifted
// by fact2_s = 2 fact_s - (T(x) - 2 T(x + s e) + T(x + 2 s e)).

// <x, M y>: one fused multiply-add per coefficient of M after the (rounded)
// coordinate products; off-diagonal pairs as fma(x_i, y_j, x_j y_i).
template <class S, int D, bool DIAG>
__device__ inline S factor_dot(const S* m, const S* x, const S* y) {
#pragma clang fp contract(off)
  S r = S(0);
  if constexpr (DIAG) {
#pragma unroll
    for (int i = 0; i < D; ++i) r = fma_(x[i] * y[i], m[i], r);
  } else {
    int k = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
      for (int j = 0; j <= i; ++j, ++k) {
        const S p = (i == j) ? x[i] * y[i] : fma_(x[i], y[j], x[j] * y[i]);
        r = fma_(p, m[k], r);
      }
    }
  }
  return r;
}

// fact[0], fact[1]: shifts of x - e and x + e; fact2 (nullable) likewise
// for x - 2e and x + 2e.  Differences of norms in cancellation-free form.
template <class S>
__device__ inline void euclid_factor(S xx, S xe, S ee, S* fact, S* fact2) {
#pragma clang fp contract(off)
  const S nx = sqrt_(xx);
  const S nxm = sqrt_(xx - S(2) * xe + ee);
  const S nxp = sqrt_(xx + S(2) * xe + ee);
  const S dm = (S(2) * xe - ee) / (nx + nxm);   // |x| - |x - e|
  const S dp = (-S(2) * xe - ee) / (nx + nxp);  // |x| - |x + e|
  const S ge = xe / nx;                         // <e, grad T(x)>
  fact[0] = dm - ge;
  fact[1] = ge + dp;
  if (fact2) {
    const S nxm2 = sqrt_(xx - S(4) * xe + S(4) * ee);
    const S nxp2 = sqrt_(xx + S(4) * xe + S(4) * ee);
    const S am = (-S(2) * xe + S(3) * ee) / (nxm + nxm2);  // |x - 2e| - |x - e|
    const S ap = (S(2) * xe + S(3) * ee) / (nxp + nxp2);   // |x + 2e| - |x + e|
    fact2[0] = S(2) * fact[0] - (dm + am);
    fact2[1] = S(2) * fact[1] - (dp + ap);
  }
}

// Adds the factorization shifts of the NS offsets e[k] of the cell at grid
// position xt to sh[2k + s] (and sets sh2, when given): the drift shift
// already in sh is kept (fact_0 + drift, fact_1 - drift as
// FiniteDifferences.h:41-45).  No-op outside the ball.
template <class S, int MODEL, int D, int NS>
__device__ inline void factor_shifts(const FactorDesc<S>& f, const int* xt, const int (&e)[NS][3],
                                     S* sh, S* sh2) {
#pragma clang fp contract(off)
  S x[3] = {S(0), S(0), S(0)};
  S n2 = S(0);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    x[c] = S(xt[c]) - f.origin[c];
    n2 = fma_(x[c], x[c], n2);
  }
  if (!(n2 < f.r2)) return;
  constexpr bool DIAG = MODEL == AGD_ISOTROPIC || MODEL == AGD_DIAGONAL;
  const S xx = factor_dot<S, D, DIAG>(f.m, x, x);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    S ef[3] = {S(e[k][0]), S(e[k][1]), S(e[k][2])};
    const S xe = factor_dot<S, D, DIAG>(f.m, x, ef);
    const S ee = factor_dot<S, D, DIAG>(f.m, ef, ef);
    S fact[2], fact2[2];
    euclid_factor<S>(xx, xe, ee, fact, sh2 ? fact2 : nullptr);
    sh[2 * k] = fact[0] + sh[2 * k];
    sh[2 * k + 1] = fact[1] + sh[2 * k + 1];
    if (sh2) {
      sh2[2 * k] = fact2[0];
      sh2[2 * k + 1] = fact2[1];
    }
  }
}

__device__ inline int32_t read_score(const int32_t* arr, int32_t tile, int32_t iter) {
  const int32_t v = arr[tile];
  return ((v >> 8) == iter) ? (v & 255) : 0;
}

// ------------------------------------------------------------------ update
// Block-wide "any": one ballot for single-wave tiles, an LDS reduction
// otherwise.
template <int NI> __device__ inline bool block_any(bool p) {
  if constexpr (NI <= 64) {
    __syncthreads();
    return __ballot(p) != 0;
  } else {
    return __syncthreads_or(p) != 0;
  }
}

// Block-wide minimum (the same value in every thread).
template <int NI, class S> __device__ inline S block_min(S v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const S o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  if constexpr (NI > 64) {
    __shared__ S part[NI / 64];
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    v = part[0];
#pragma unroll
    for (int w = 1; w < NI / 64; ++w) v = part[w] < v ? part[w] : v;
  }
  return v;
}

// Resolves the 2*nsym neighbours x +- e_k of a cell into one code each:
//   code >= 0 : tiled index of a neighbour outside the tile (gathered),
//   code <= -2: LDS slot (-2 - code) inside the tile,
//   code == -1: no neighbour (outside the domain or hidden by a wall).
template <class S, int MODEL, int D, class TL, bool GENERAL>
__device__ inline void resolve_codes(const Stencil<S, MODEL, D>& st, const int* xi, const int* xt,
                                     const int* xo, int wx, const uint8_t* wall_s,
                                     const uint8_t* wall, const GridDesc& g, int* code,
                                     int tile_base = 0, int* code2 = nullptr) {
  constexpr int NS = ModelTraits<MODEL, D>::NSYM;
  constexpr int NI = TL::NI;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int j = 2 * k + s;
      const int sg = 2 * s - 1;
      int e[3] = {sg * st.e[k][0], sg * st.e[k][1], sg * st.e[k][2]};
      if constexpr (!GENERAL) {
        // No walls, no periodic axes: unsigned range checks, tile
        // coordinates by shifts, tile-local test against the tile's own
        // (uniform) coordinates.
        int yt[3] = {xt[0] + e[0], xt[1] + e[1], xt[2] + e[2]};
        bool dom = true;
#pragma unroll
        for (int c = 0; c < D; ++c) dom = dom && unsigned(yt[c]) < unsigned(g.shape_tot[c]);
        int yo[3] = {0, 0, 0}, yi[3] = {0, 0, 0};
        bool local = true;
#pragma unroll
        for (int c = 0; c < D; ++c) {
          const unsigned u = unsigned(yt[c]);
          yo[c] = int(u / unsigned(TL::T[c]));
          yi[c] = int(u % unsigned(TL::T[c]));
          local = local && yo[c] == xo[c];
        }
        const int li = TL::index(yi);
        int to = yo[0];
#pragma unroll
        for (int c = 1; c < D; ++c) to = imul24(to, g.shape_o[c]) + yo[c];
        code[j] = !dom ? -1 : (local ? -2 - li : (tile_base + to) * NI + li);
      } else {
        bool vis = true;
        if (wall && wx != 255) {
          vis = visible<D>(e, wx, [&](const int* w) -> int {
            int yi[3] = {xi[0] + w[0], xi[1] + w[1], xi[2] + w[2]};
            int yt[3] = {xt[0] + w[0], xt[1] + w[1], xt[2] + w[2]};
            bool dom = true;
#pragma unroll
            for (int c = 0; c < D; ++c)
              dom = dom && (((g.periodic >> c) & 1) || (yt[c] >= 0 && yt[c] < g.shape_tot[c]));
            if (TL::inside(yi) && dom) return wall_s[TL::index(yi)];
            const int m = index_tot<D, TL>(yt, g);
            return m >= 0 ? int(wall[m]) : 255;
          });
        }
        int yi[3] = {xi[0] + e[0], xi[1] + e[1], xi[2] + e[2]};
        int yt[3] = {xt[0] + e[0], xt[1] + e[1], xt[2] + e[2]};
        bool in_dom = true;
#pragma unroll
        for (int c = 0; c < D; ++c) in_dom = in_dom && (yt[c] >= 0 && yt[c] < g.shape_tot[c]);
        const bool local = vis && TL::inside(yi) && in_dom;
        const int m = (vis && !local) ? index_tot<D, TL>(yt, g) : -1;
        code[j] = local ? -2 - TL::index(yi) : m;
        if (code2) {
          // further neighbour x + 2 sg e on the same side, behind the same
          // visibility test (FiniteDifferences.h:75-90)
          int y2i[3] = {yi[0] + e[0], yi[1] + e[1], yi[2] + e[2]};
          int y2t[3] = {yt[0] + e[0], yt[1] + e[1], yt[2] + e[2]};
          bool in2 = true;
#pragma unroll
          for (int c = 0; c < D; ++c) in2 = in2 && (y2t[c] >= 0 && y2t[c] < g.shape_tot[c]);
          const bool local2 = vis && TL::inside(y2i) && in2;
          const int m2 = (vis && !local2) ? index_tot<D, TL>(y2t, g) : -1;
          code2[j] = local2 ? -2 - TL::index(y2i) : m2;
        }
      }
    }
  }
}

// Kernel variants: VAR_FAST assumes no walls and no periodic axes, so its
// neighbour indexing carries no wrap-around code; VAR_GENERAL handles both;
// VAR_PRE reads stencil weights and neighbour codes precomputed once per
// solve (any walls / periodicity are already resolved in the codes);
// VAR_BATCH is VAR_FAST over the concatenated tiles of several instances,
// each tile looking up its instance's grid and constants; VAR_ORDER2 is
// VAR_GENERAL with the second-order correction (HFM.h:91-134).
enum { VAR_FAST = 0, VAR_GENERAL = 1, VAR_PRE = 2, VAR_BATCH = 3, VAR_ORDER2 = 4 };

// Tiles of fewer than 64 cells (2D 4x4, 4x8, 8x4; batched solves only) are
// packed K = 64 / NI to a single-wave workgroup: lane group g = tid / NI
// visits list entry (workgroup's first entry + g) with its own LDS slice, its
// convergence flag a ballot restricted to its lanes, and its activation
// lanes g * NI + (0 .. 2D).
template <int NI> constexpr int tiles_per_block() { return NI < 64 ? 64 / NI : 1; }
template <int NI> constexpr int block_lanes() { return NI < 64 ? 64 : NI; }

// (Elastica2 asks for two resident waves per SIMD: its 30 offsets per cell
// otherwise hold 326 registers and one wave; measured 42.6 -> 37.4 ms on
// 192 x 192 x 64 despite 300 bytes of spills per lane)
template <class S, int MODEL, int D, int T0, int T1, int T2, int VAR>
__global__ void __launch_bounds__(block_lanes<T0 * T1 * T2>(), MODEL == AGD_ELASTICA2 ? 2 : 1)
    eik_update(EikArgs<S> a) {
  constexpr bool ORDER2 = VAR == VAR_ORDER2;
  constexpr bool GENERAL = VAR == VAR_GENERAL || ORDER2;
  constexpr bool WALLS = GENERAL;
  constexpr bool BATCH = VAR == VAR_BATCH;
  // Every variant but the general ones addresses per-cell arrays with 32-bit
  // byte offsets (the host sends grids of 4 GiB per field or more to the
  // general variant).
  constexpr bool OFF32 = !GENERAL;
  auto ldu = [&](const S* base, int idx) -> S {
    if constexpr (OFF32) return ld32(base, uint32_t(idx));
    else return base[idx];
  };
  using TL = Tile<D, T0, T1, T2>;
  using MT = ModelTraits<MODEL, D>;
  constexpr int NI = TL::NI;
  constexpr int NS = MT::NSYM;
  constexpr int NG = MT::NGEOM;
  constexpr int NJ = 2 * NS;
  constexpr int K = tiles_per_block<NI>();
  static_assert(NI % 64 == 0 || (BATCH && 64 % NI == 0),
                "tiles are whole waves, or (batched solves) pack whole into one wave");
  // tile g's values at [g * (NI + 1), g * (NI + 1) + NI); slot NI holds 0:
  // the read of neighbours outside the tile
  __shared__ S u_s[K * (NI + 1)];
  __shared__ uint8_t wall_s[WALLS ? NI : 1];

  const int tid = threadIdx.x;
  const int lt = K > 1 ? tid % NI : tid;  // cell of the tile
  const int grp = K > 1 ? tid / NI : 0;   // tile of the workgroup
  S* const us = u_s + grp * (NI + 1);
  const PassView pv = pass_begin(a.w);
  const int count = pv.count;
  int xi[3];
  TL::pos(lt, xi);
  const int32_t tag = pv.iter + 1;
  const bool fim = a.solver == AGD_FIM;
  // bound_active_blocks (general variants only): the threshold of this pass
  // (set after the previous one)
  const bool bab = GENERAL && a.bab.ctl != nullptr;
  const S bab_thr = bab ? S(reinterpret_cast<const double*>(a.bab.ctl + kBabThr)[1]) : inf_v<S>();

  for (int li = blockIdx.x * K; li < count; li += gridDim.x * K) {
    // A lane group past the end of the list repeats the last entry with
    // every store and activation switched off.
    int pk;
    bool valid = true;
    if constexpr (K == 1) {
      pk = pass_entry(pv, li);
    } else {
      pk = 0;
#pragma unroll
      for (int g = 0; g < K; ++g) {
        const int e = pass_entry(pv, min(li + g, count - 1));
        if (g == grp) pk = e;
      }
      valid = li + grp < count;
    }
    int xo[3];
    // Grid and constants of the tile's instance (the solve's own when not
    // batched).  Batched list entries are tile ids, else packed coordinates.
    GridDesc gi;
    int tbase = 0;
    S atol, rtol, iso_w[3];
    int n_o;
    if constexpr (BATCH) {
      n_o = K == 1 ? __builtin_amdgcn_readfirstlane(pk) : pk;
      const int ii = K == 1 ? __builtin_amdgcn_readfirstlane(a.tile_inst[n_o]) : a.tile_inst[n_o];
      const InstDesc* I = a.inst + ii;
      gi = I->g;
      tbase = I->tile_base;
      atol = S(I->atol);
      rtol = S(I->rtol);
#pragma unroll
      for (int k = 0; k < 3; ++k) iso_w[k] = S(I->iso_w[k]);
      tile_pos<D>(n_o - tbase, gi, xo);
    } else {
      gi = a.g;
      atol = a.atol;
      rtol = a.rtol;
#pragma unroll
      for (int k = 0; k < 3; ++k) iso_w[k] = a.iso_w[k];
      unpack_tile<D>(pk, xo);
      n_o = tile_linear<D>(xo, a.g);
    }
    int xt[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) xt[k] = xo[k] * TL::T[k] + xi[k];
    const int n_t = n_o * NI + lt;

    // ---- activation inputs first: the tile and its face neighbours
    // (lanes 0..2D of the tile's lanes), and their FIM scores of the
    // previous pass.
    int nb = -1, nbp = pk, sc = 0;
    if (lt < 2 * D + 1) {
      if (lt == 2 * D) {
        nb = n_o;
      } else {
        nb = tile_neighbor<D>(xo, lt >> 1, 2 * (lt & 1) - 1, gi, &nbp);
        if (BATCH && nb >= 0) {
          nb += tbase;
          nbp = nb;
        }
      }
      if (fim && nb >= 0) sc = read_score(pv.score_prev, nb, pv.iter - 1);
    }

    // ---- bound_active_blocks: a tile listed only because it is frozen
    // (no update was requested for this pass) skips the update and only
    // re-examines its pending change value against the threshold.
    bool real = true;
    S pend = inf_v<S>();
    if (bab) {
      real = __builtin_amdgcn_readfirstlane(a.bab.needs[n_o]) >= pv.iter;
      pend = S(a.bab.pending[n_o]);
    }
    bool any = false;
    S minchg = inf_v<S>();
    if (real) {
    // ---- cell loads (coalesced: a tile is contiguous)
    const S u_old = ldu(a.u, n_t);
    const S rhs = ldu(a.rhs, n_t);
    const bool seed = (a.seeds[n_t >> 5] >> (n_t & 31)) & 1u;
    int wx = 255;
    if constexpr (WALLS) {
      if (a.wall) wx = a.wall[n_t];
      wall_s[tid] = uint8_t(wx);
    }
    us[lt] = seed ? rhs : u_old;
    if (lt == 0) us[NI] = S(0);

    int code[NJ];
    int code2[ORDER2 ? NJ : 1];
    // sh[2k + s]: shift of the value of neighbour x -+ e_k (drift, source
    // factorization); sh2 likewise for the further neighbours (order 2).
    S sw[NS], sh[NJ];
    S sh2[ORDER2 ? NJ : 1];
    if constexpr (ORDER2) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) sh2[j] = S(0);
    }
    if constexpr (MT::STORED) {
      // stored sub-schemes: read per sweep below
    } else if constexpr (VAR == VAR_PRE) {
      // Stencil weights and neighbour codes precomputed for the whole solve.
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        sw[k] = ld32(a.sw + k * a.g.size_tot, uint32_t(n_t));
        sh[2 * k] = S(0);
        sh[2 * k + 1] = S(0);
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) code[j] = ld32(a.scode + j * a.g.size_tot, uint32_t(n_t));
    } else {
      S geo[NG > 0 ? NG : 1];
#pragma unroll
      for (int k = 0; k < NG; ++k) {
        const S* __restrict__ gk = a.geom + k * a.g.size_tot;  // uniform base
        geo[k] = ldu(gk, n_t);
      }
      Stencil<S, MODEL, D> st;
      make_stencil<S, MODEL, D>(geo, iso_w, st);
      if constexpr (WALLS) __syncthreads();  // wall_s complete before the visibility walks
      resolve_codes<S, MODEL, D, TL, GENERAL>(st, xi, xt, xo, wx, wall_s, a.wall, gi, code, tbase,
                                              ORDER2 ? code2 : nullptr);
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        sw[k] = st.w[k];
        sh[2 * k] = st.shift[k];
        sh[2 * k + 1] = -st.shift[k];
      }
      if constexpr (GENERAL) {
        if (a.fac.r2 > S(0)) factor_shifts<S, MODEL, D, NS>(a.fac, xt, st.e, sh, ORDER2 ? sh2 : nullptr);
      }
    }
    if constexpr (MT::STORED) {
      // ---- TTI and the curvature models: the min or max over the
      // sub-schemes of their upwind updates (HFMIter.h:29-56, strict in-tile
      // iteration).  Weights and neighbour codes are re-read at every sweep
      // (cache hits) rather than held in registers; out-of-tile values are
      // read at every sweep.  A forward offset's x - e code is -1.
      const int64_t N = a.g.size_tot;
      const bool mix_min = ld32(a.scode + int64_t(MT::NMIX * NJ) * N, uint32_t(n_t)) != 0;
      // source factorization: the cell's row of the shift table, or -1
      const int64_t fb = a.fac.r2 > S(0) ? factor_box_index<S, D>(a.fac, xt) : -1;
      const int64_t nbox = a.fac.nbox;
      const int64_t fb2 = fb + int64_t(MT::NMIX * NJ) * nbox;  // further neighbours
      __syncthreads();  // u_s complete
      const bool active = !seed;
      for (int it = 0; it < a.niter_i; ++it) {
        S best = mix_min ? inf_v<S>() : -inf_v<S>();
#pragma unroll 1
        for (int m = 0; m < MT::NMIX; ++m) {
          S v[NS], w[NS], v2[MT::CURV ? 1 : NS];
#pragma unroll
          for (int k = 0; k < NS; ++k) {
            w[k] = ld32(a.sw + int64_t(m * NS + k) * N, uint32_t(n_t));
            S side[2];
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
              const int c = ld32(a.scode + int64_t(m * NJ + 2 * k + sd) * N, uint32_t(n_t));
              side[sd] = c <= -2 ? us[-2 - c] : (c >= 0 ? ld32(a.u, uint32_t(c)) : inf_v<S>());
              // (the curvature models have no source factorization)
              if constexpr (!MT::CURV)
                if (fb >= 0) side[sd] += a.fac.tab[int64_t(m * NJ + 2 * k + sd) * nbox + fb];
            }
            const int sd = (side[0] > side[1]) ? 1 : 0;  // Greater(v,-) keeps the -e side on ties
            v[k] = side[sd];
            // curvature models: an offset of weight 0 (pruned, or a Convex
            // variant's right turn) takes no part in the update; left in, a
            // much smaller value of it would set the base of the quadratic
            // and cancel its FP32 digits
            if constexpr (MT::CURV) v[k] = w[k] > S(0) ? v[k] : inf_v<S>();
            if constexpr (!MT::CURV) {  // (the curvature models are first order)
              if (a.scode2) {
                // the further neighbour on the chosen side (second order)
                const int c2 = ld32(a.scode2 + int64_t(m * NJ + 2 * k + sd) * N, uint32_t(n_t));
                v2[k] = c2 <= -2 ? us[-2 - c2] : (c2 >= 0 ? ld32(a.u, uint32_t(c2)) : inf_v<S>());
                if (fb >= 0) v2[k] += a.fac.tab[int64_t(m * NJ + 2 * k + sd) * nbox + fb2];
              }
            }
          }
          if constexpr (!MT::CURV)
            if (a.scode2) order2_correct<S, NS>(us[lt], a.order2_thr, v, v2, w);
          const S um = hfm_solve<S, NS>(v, w, rhs);
          if (mix_min == (best > um)) best = um;
        }
        __syncthreads();
        const bool ch = active && (us[lt] > best);
        if (ch) us[lt] = best;
        if (!block_any<NI>(ch)) break;
      }
    } else {
      // All gathers are issued before any is consumed (safe address = own
      // cell), so a visit costs a single memory round trip.
      S g[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) g[j] = ldu(a.u, code[j] >= 0 ? code[j] : n_t);
      S g2[ORDER2 ? NJ : 1];
      if constexpr (ORDER2) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) g2[j] = a.u[code2[j] >= 0 ? code2[j] : n_t];
      }
      S vo[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) vo[j] = code[j] <= -2 ? sh[j] : (code[j] >= 0 ? g[j] + sh[j] : inf_v<S>());
      // LDS slot read at every sweep: the neighbour's own slot inside the
      // tile, slot NI (holding 0) otherwise, whose value vo then passes
      // through unchanged (0 + vo == vo up to the sign of a zero, which no
      // comparison or later sum distinguishes).  The sweep carries no
      // branch per neighbour.
      int slot[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) slot[j] = code[j] <= -2 ? -2 - code[j] : NI;
      __syncthreads();  // u_s complete

      // ---- in-tile sweeps (HFMIter.h:25-84, DECREASING)
      const bool active = !seed;
      for (int it = 0; it < a.niter_i; ++it) {
        S v[NS], w[NS];
        S w2v[ORDER2 ? NS : 1];  // further-neighbour values (order 2)
        S x[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) x[j] = us[slot[j]];
        // Keeps every read unconditional (the compiler would otherwise sink
        // each into a branch with its own wait on the LDS).
#pragma unroll
        for (int j = 0; j < NJ; ++j) asm volatile("" ::"v"(x[j]));
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          const S a0 = x[2 * k] + vo[2 * k];
          const S a1 = x[2 * k + 1] + vo[2 * k + 1];
          v[k] = (a0 > a1) ? a1 : a0;  // Greater(v,-) keeps the -e side on ties
          w[k] = sw[k];
          if constexpr (ORDER2) {
            // the further neighbour on the chosen side
            const int j2 = 2 * k + ((a0 > a1) ? 1 : 0);
            const int c2 = code2[j2];
            w2v[k] = c2 <= -2 ? us[-2 - c2] + sh2[j2] : (c2 >= 0 ? g2[j2] + sh2[j2] : inf_v<S>());
          }
        }
        if constexpr (ORDER2) order2_correct<S, NS>(us[lt], a.order2_thr, v, w2v, w);
        const S un = hfm_solve<S, NS>(v, w, rhs);
        __syncthreads();
        const bool ch = active && (us[lt] > un);
        if (ch) us[lt] = un;
        // (packed tiles: the sweeps go on while any tile of the wave
        // changes; a tile that did not change gives the same values again)
        if (!block_any<NI>(ch)) break;
      }
    }

    // ---- store and convergence (Update.h:163-219)
    const S un = us[lt];
    if (valid) {
      if constexpr (OFF32) st32(a.u, uint32_t(n_t), un);
      else a.u[n_t] = un;
    }
    const S diff = abs_(u_old - un);
    const S tol = atol + rtol * abs_(un);
    const bool changed = !(isnan_(diff) || diff <= tol);
    if constexpr (K == 1) {
      any = block_any<NI>(changed);
    } else {
      __syncthreads();
      const uint64_t m = (uint64_t(1) << NI) - 1;
      any = ((uint64_t(__ballot(changed)) >> (grp * NI)) & m) != 0;
    }
    if (bab) minchg = block_min<NI>(changed ? un : inf_v<S>());
    }  // real

    // ---- activation (Propagation.h:103-167), wave 0 only.  The
    // global-iteration solver runs with AGSI tags: they only decide when to
    // stop, every pass visits the full list.
    if (tid < 64) {
      if (bab) {
        // Propagation.h:116-126 with the change still pending from a frozen
        // visit: propagate (self and faces, with an update request) below
        // the threshold, else stay frozen (self only) while a change waits.
        const S m = minchg < pend ? minchg : pend;
        const bool prop = m < bab_thr;
        const bool frz = !prop && m != inf_v<S>();
        if (prop && nb >= 0) atomicMax(a.bab.needs + nb, tag);
        const bool want = prop ? nb >= 0 : (frz && tid == 2 * D);
        enqueue_wave(want, nb, nbp, tag, a.w.stamp, pv.list_out, pv.count_out);
        if (tid == 0) {
          a.bab.pending[n_o] = frz ? double(m) : __builtin_huge_val();
          const int sh = (blockIdx.x % kShards) * kCountStride;
          if (prop) atomicAdd(a.bab.ctl + kBabProp + sh, 1);
          if (frz) atomicAdd(a.bab.ctl + kBabFrozen + sh, 1);
          if (m != inf_v<S>()) {
            atomicMin(a.bab.ctl + kBabMin + sh, ord_int(float(m)));
            atomicMax(a.bab.ctl + kBabMax + sh, ord_int(float(m)));
          }
        }
      } else if (!fim) {
        enqueue_wave(valid && any && nb >= 0, nb, nbp, tag, a.w.stamp, pv.list_out, pv.count_out);
      } else {
        const int front = a.fim_front;
        const int l0 = grp * NI;  // the tile's first lane
        const int prev = __shfl(sc, l0 + 2 * D);
        int mx = 0;
#pragma unroll
        for (int i = 0; i < 2 * D; ++i) mx = max(mx, __shfl(sc, l0 + i));
        int next;
        bool prop;
        if (!any) {
          if (prev == 0 || prev == front) { next = 0; prop = false; }
          else { next = front; prop = true; }
        } else {
          next = 1;
          if (front > 2) next = max(next, mx - 1);
          prop = next > 1;
        }
        if (lt == 2 * D && valid) pv.score_next[n_o] = (pv.iter << 8) | next;
        const bool want = valid && ((lt == 2 * D) ? any : (prop && lt < 2 * D && nb >= 0));
        enqueue_wave(want, nb, nbp, tag, a.w.stamp, pv.list_out, pv.count_out);
      }
    }
    __syncthreads();  // LDS reuse by the next tile of this workgroup
  }
}

// ------------------------------------------------------------------ scheme
// Stencil weights and neighbour codes of every cell, computed once per solve
// for the VAR_PRE update (the counterpart of the reference's
// precompute_scheme export kernel, _Kernel.py:254-268 and 346-359, with
// offsets stored already resolved against the tiled grid).
template <class S> struct SchemeArgs {
  const S* geom;
  const uint8_t* wall;
  GridDesc g;
  S iso_w[3];
  S* sw;
  int32_t* scode;
  int32_t* scode2;  // TTI, second order: codes of the further neighbours x -+ 2 e_k, or null
  S relax;          // curvature models: relaxation of the outer products (eps^2)
  int32_t convex;   // curvature models: the Convex variant
};

template <class S, int MODEL, int D, int T0, int T1, int T2, bool GENERAL>
__global__ void __launch_bounds__(T0* T1* T2) eik_scheme(SchemeArgs<S> a) {
  using TL = Tile<D, T0, T1, T2>;
  using MT = ModelTraits<MODEL, D>;
  constexpr int NI = TL::NI;
  constexpr int NS = MT::NSYM;
  constexpr int NG = MT::NGEOM;
  __shared__ uint8_t wall_s[GENERAL ? NI : 1];
  const int tid = threadIdx.x;
  const int n_o = blockIdx.x;
  int xi[3], xo[3], xt[3] = {0, 0, 0};
  TL::pos(tid, xi);
  tile_pos<D>(n_o, a.g, xo);
#pragma unroll
  for (int k = 0; k < D; ++k) xt[k] = xo[k] * TL::T[k] + xi[k];
  const int64_t N = a.g.size_tot;
  const int64_t n_t = int64_t(n_o) * NI + tid;
  int wx = 255;
  if constexpr (GENERAL) {
    if (a.wall) wx = a.wall[n_t];
    wall_s[tid] = uint8_t(wx);
    __syncthreads();
  }
  S geo[NG > 0 ? NG : 1];
#pragma unroll
  for (int k = 0; k < NG; ++k) geo[k] = a.geom[k * N + n_t];
  if constexpr (MT::CURV) {
    // Curvature models: the stencils of every sub-scheme at once, forward
    // slots with their x - e code set to -1, then the min/max flag.
    S cw[MT::NMIX * NS];
    int ce[MT::NMIX * NS][3];
    const bool mix_min = curv_stencil<S, MODEL>(geo, a.relax, a.convex != 0, cw, ce);
#pragma unroll 1
    for (int m = 0; m < MT::NMIX; ++m) {
      Stencil<S, MODEL, D> st;
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        st.w[k] = cw[m * NS + k];
#pragma unroll
        for (int c = 0; c < 3; ++c) st.e[k][c] = ce[m * NS + k][c];
        st.shift[k] = S(0);
      }
      int code[2 * NS];
      resolve_codes<S, MODEL, D, TL, true>(st, xi, xt, xo, wx, wall_s, a.wall, a.g, code, 0, nullptr);
#pragma unroll
      for (int k = NS - MT::NFWD; k < NS; ++k) code[2 * k] = -1;
      if (MODEL == AGD_REEDSSHEPPFWD2 && a.convex) code[0] = -1;  // rotation to the left only
#pragma unroll
      for (int k = 0; k < NS; ++k) a.sw[int64_t(m * NS + k) * N + n_t] = st.w[k];
#pragma unroll
      for (int j = 0; j < 2 * NS; ++j) a.scode[int64_t(m * 2 * NS + j) * N + n_t] = code[j];
    }
    a.scode[int64_t(MT::NMIX * 2 * NS) * N + n_t] = mix_min ? 1 : 0;
  } else if constexpr (MT::NMIX > 1) {
    // TTI_.h:112-134: one Selling decomposition per tangent ellipse, then
    // the min/max flag after the NMIX*2*NS codes.
    S diag[kTTIMix][2];
    const bool mix_min = tti_diags<S>(geo, geo + 2, diag);
    S D0[NS], D1[NS];
    tti_base<S, D>(geo, D0, D1);
#pragma unroll 1
    for (int m = 0; m < MT::NMIX; ++m) {
      Stencil<S, MODEL, D> st;
      tti_sub<S, D>(D0, D1, diag[m][0], diag[m][1], st.w, st.e);
      int code[2 * NS];
      int code2[2 * NS];
      resolve_codes<S, MODEL, D, TL, GENERAL>(st, xi, xt, xo, wx, wall_s, a.wall, a.g, code, 0,
                                              (GENERAL && a.scode2) ? code2 : nullptr);
      if (!GENERAL && a.scode2) {
        // no walls: the further neighbours are the codes of the doubled offsets
        Stencil<S, MODEL, D> st2 = st;
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
          for (int c = 0; c < 3; ++c) st2.e[k][c] = 2 * st.e[k][c];
        resolve_codes<S, MODEL, D, TL, GENERAL>(st2, xi, xt, xo, wx, wall_s, a.wall, a.g, code2);
      }
#pragma unroll
      for (int k = 0; k < NS; ++k) a.sw[int64_t(m * NS + k) * N + n_t] = st.w[k];
#pragma unroll
      for (int j = 0; j < 2 * NS; ++j) a.scode[int64_t(m * 2 * NS + j) * N + n_t] = code[j];
      if (a.scode2) {
#pragma unroll
        for (int j = 0; j < 2 * NS; ++j) a.scode2[int64_t(m * 2 * NS + j) * N + n_t] = code2[j];
      }
    }
    a.scode[int64_t(MT::NMIX * 2 * NS) * N + n_t] = mix_min ? 1 : 0;
  } else {
    Stencil<S, MODEL, D> st;
    make_stencil<S, MODEL, D>(geo, a.iso_w, st);
    int code[2 * NS];
    resolve_codes<S, MODEL, D, TL, GENERAL>(st, xi, xt, xo, wx, wall_s, a.wall, a.g, code);
#pragma unroll
    for (int k = 0; k < NS; ++k) a.sw[k * N + n_t] = st.w[k];
#pragma unroll
    for (int j = 0; j < 2 * NS; ++j) a.scode[j * N + n_t] = code[j];
  }
}

// ------------------------------------------------------------------ TTI factor
// The source-factorization shifts of every TTI sub-scheme at the cells of
// the factoring ball (FiniteDifferences.h:30-45 with TTI_.h:191-211), written
// into the table of FactorDesc once per solve.  The sub-stencils are rebuilt
// from the geometry as eik_scheme builds them, so that entry (m, k) belongs to
// the offset whose codes the scheme stored.
template <class S> struct TTIFactorArgs {
  const S* geom;
  GridDesc g;
  FactorDesc<S> fac;
  S* tab;
  int32_t order2;  // also the shifts of the further neighbours
};

template <class S, int D, int T0, int T1, int T2>
__global__ void __launch_bounds__(T0* T1* T2) eik_tti_factor(TTIFactorArgs<S> a) {
  using TL = Tile<D, T0, T1, T2>;
  using MT = ModelTraits<AGD_TTI, D>;
  constexpr int NI = TL::NI;
  constexpr int NS = MT::NSYM;
  constexpr int NG = MT::NGEOM;
  const int tid = threadIdx.x;
  const int n_o = blockIdx.x;
  int xi[3], xo[3], xt[3] = {0, 0, 0};
  TL::pos(tid, xi);
  tile_pos<D>(n_o, a.g, xo);
#pragma unroll
  for (int k = 0; k < D; ++k) xt[k] = xo[k] * TL::T[k] + xi[k];
  const int64_t b = factor_box_index<S, D>(a.fac, xt);
  if (b < 0) return;
  const int64_t N = a.g.size_tot;
  const int64_t n_t = int64_t(n_o) * NI + tid;
  S geo[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) geo[k] = a.geom[k * N + n_t];
  S diag[kTTIMix][2];
  (void)tti_diags<S>(geo, geo + 2, diag);
  S D0[NS], D1[NS];
  tti_base<S, D>(geo, D0, D1);
  TTIFactorData<S> t;
  tti_factor_data<S>(a.fac.m, t);
  S x[D], grad[D];
  {
#pragma clang fp contract(off)
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = S(xt[c]) - a.fac.origin[c];
  }
  const S nx = tti_factor_norm<S, D>(a.fac.m, t, x, grad);
  const int64_t nb = a.fac.nbox;
  const int64_t off2 = int64_t(kTTIMix) * 2 * NS * nb;
#pragma unroll 1
  for (int m = 0; m < kTTIMix; ++m) {
    S w[NS];
    int e[NS][3];
    tti_sub<S, D>(D0, D1, diag[m][0], diag[m][1], w, e);
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      S f[2], f2[2];
      tti_factor_pair<S, D>(a.fac.m, t, x, nx, grad, e[k], f, a.order2 ? f2 : nullptr);
      const int64_t c = int64_t(m * 2 * NS + 2 * k);
      a.tab[c * nb + b] = f[0];
      a.tab[(c + 1) * nb + b] = f[1];
      if (a.order2) {
        a.tab[off2 + c * nb + b] = f2[0];
        a.tab[off2 + (c + 1) * nb + b] = f2[1];
      }
    }
  }
}

// ------------------------------------------------------------------ flow
// One global pass with niter_i = 1 and flow export (Update.h:149-209).
template <class S> struct FlowArgs {
  S* u;
  const S* geom;
  const S* rhs;
  const uint32_t* seeds;
  const uint8_t* wall;
  GridDesc g;
  S iso_w[3];
  S* fw;
  S* wsum;
  int32_t* fidx;
  S* fvec;
  int32_t order2;  // second-order correction
  S order2_thr;
  // Residual check instead of export: count the non-seed cells whose update
  // exceeds their value by more than chk_atol + chk_rtol |u| (a value below
  // the fixed point, which the decreasing iteration cannot raise).
  int32_t* violations;
  S chk_atol, chk_rtol;
  // With violations: 1 also raises every counted cell to +inf (a value at
  // +inf is above the fixed point; repeated until nothing is counted, this
  // turns a start into a supersolution of the scheme).
  int32_t repair;
  // stale: count instead the cells whose update lies below their value by
  // more than the tolerance (a value the solve left above the fixed point),
  // and set their bits in stale_bits when given.
  int32_t stale;
  uint32_t* stale_bits;
  FactorDesc<S> fac;
  S relax;  // curvature models: relaxation of the outer products (eps^2)
  int32_t convex;  // curvature models: the Convex variant
};

// The upwind update of one cell and its flow (Update.h:149-209, the body
// of the flow pass): value val of the scheme at the cell, flow weights
// w_k (v - v_k)_+ (3/2 of it where the second-order correction applied),
// their sum, the flow vector sum_k fw_k (side_k e_k) in grid units and the
// chosen neighbours' tiled indices.  Reads the cell and its neighbours only
// (the flow pass and the geodesic ODE's recomputed flow share it).
template <class S, int NS> struct CellFlow {
  S u_old, val;
  bool seed;
  S fw[NS];
  int64_t nb[NS];
  S sum;
  S vec[3];
};

template <class S, int MODEL, int D, class TL>
__device__ inline void flow_cell(const FlowArgs<S>& a, const int* xt, int64_t n_t,
                                 CellFlow<S, ModelTraits<MODEL, D>::NSYM>& out) {
  using MT = ModelTraits<MODEL, D>;
  constexpr int NS = MT::NSYM;
  constexpr int NG = MT::NGEOM;
  const int64_t N = a.g.size_tot;

  const S u_old = a.u[n_t];
  const S rhs = a.rhs[n_t];
  const bool seed = (a.seeds[n_t >> 5] >> (n_t & 31)) & 1u;
  S geo[NG > 0 ? NG : 1];
#pragma unroll
  for (int k = 0; k < NG; ++k) geo[k] = a.geom[k * N + n_t];
  const int wx = a.wall ? int(a.wall[n_t]) : 255;

  // TTI source factorization: N(x) and its gradient, once per cell.
  bool tfac = false;
  TTIFactorData<S> tdat;
  S xrel[D], tgrad[D], tnx = S(0);
  if constexpr (MODEL == AGD_TTI) {
    tfac = a.fac.r2 > S(0) && factor_box_index<S, D>(a.fac, xt) >= 0;
    if (tfac) {
#pragma clang fp contract(off)
      tti_factor_data<S>(a.fac.m, tdat);
#pragma unroll
      for (int c = 0; c < D; ++c) xrel[c] = S(xt[c]) - a.fac.origin[c];
      tnx = tti_factor_norm<S, D>(a.fac.m, tdat, xrel, tgrad);
    }
  }

  // Neighbour values (min over -e/+e, ties keep -e), sides and indices of a
  // stencil.
  auto gather = [&](const Stencil<S, MODEL, D>& st, S* v, S* w, int* side, int64_t* nb,
                    S* v2) {
    S sh[2 * NS], sh2[2 * NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      sh[2 * k] = st.shift[k];
      sh[2 * k + 1] = -st.shift[k];
      sh2[2 * k] = sh2[2 * k + 1] = S(0);
    }
    if constexpr (!MT::STORED) {
      if (a.fac.r2 > S(0)) factor_shifts<S, MODEL, D, NS>(a.fac, xt, st.e, sh, v2 ? sh2 : nullptr);
    } else if constexpr (MODEL == AGD_TTI) {
      if (tfac) {
#pragma unroll 1
        for (int k = 0; k < NS; ++k)
          tti_factor_pair<S, D>(a.fac.m, tdat, xrel, tnx, tgrad, st.e[k], sh + 2 * k, v2 ? sh2 + 2 * k : nullptr);
      }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      S best = inf_v<S>(), best2 = inf_v<S>();
      int bs = 0;
      int64_t bi = -1;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int sg = 2 * s - 1;
        int e[3] = {sg * st.e[k][0], sg * st.e[k][1], sg * st.e[k][2]};
        // a forward offset (the last NFWD slots) has no x - e neighbour
        const bool none = (MT::NFWD > 0 && k >= NS - MT::NFWD && s == 0) ||
                          (MODEL == AGD_REEDSSHEPPFWD2 && a.convex && k == 0 && s == 0);
        bool vis = !none;
        if (vis && a.wall && wx != 255) {
          vis = visible<D>(e, wx, [&](const int* ww) -> int {
            int yt[3] = {xt[0] + ww[0], xt[1] + ww[1], xt[2] + ww[2]};
            const int64_t m = index_tot<D, TL>(yt, a.g);
            return m >= 0 ? int(a.wall[m]) : 255;
          });
        }
        int yt[3] = {xt[0] + e[0], xt[1] + e[1], xt[2] + e[2]};
        const int64_t m = vis ? index_tot<D, TL>(yt, a.g) : -1;
        const S val = (m >= 0) ? a.u[m] + sh[2 * k + s] : inf_v<S>();
        S val2 = inf_v<S>();
        if (v2 && vis) {
          int y2[3] = {yt[0] + e[0], yt[1] + e[1], yt[2] + e[2]};
          const int64_t m2 = index_tot<D, TL>(y2, a.g);
          if (m2 >= 0) val2 = a.u[m2] + sh2[2 * k + s];
        }
        if (s == 0) { best = val; best2 = val2; bs = 0; bi = m; }
        else if (best > val) { best = val; best2 = val2; bs = 1; bi = m; }
      }
      v[k] = best;
      if (v2) v2[k] = best2;
      side[k] = bs;
      nb[k] = bi;
      w[k] = st.w[k];
    }
  };

  Stencil<S, MODEL, D> st;  // the (chosen sub-)stencil
  S v[NS], w[NS];
  int side[NS];
  int64_t nb[NS];
  S val;
  unsigned o2mask = 0;  // order 2: neighbours with the second-order correction
  if constexpr (MT::STORED) {
    // TTI and the curvature models: keep the sub-scheme whose update wins
    // the min/max (HFMIter.h:48-54).
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      v[k] = inf_v<S>();
      w[k] = S(0);
      side[k] = 0;
      nb[k] = -1;
      st.w[k] = S(0);
      st.shift[k] = S(0);
      st.e[k][0] = st.e[k][1] = st.e[k][2] = 0;
    }
    bool mix_min;
    S diag[MT::CURV ? 1 : kTTIMix][2];
    S D0[NS], D1[NS];
    S cw[MT::CURV ? MT::NMIX * NS : 1];
    int ce[MT::CURV ? MT::NMIX * NS : 1][3];
    if constexpr (MT::CURV) {
      mix_min = curv_stencil<S, MODEL>(geo, a.relax, a.convex != 0, cw, ce);
    } else {
      mix_min = tti_diags<S>(geo, geo + 2, diag);
      tti_base<S, D>(geo, D0, D1);
    }
    val = mix_min ? inf_v<S>() : -inf_v<S>();
#pragma unroll 1
    for (int m = 0; m < MT::NMIX; ++m) {
      Stencil<S, MODEL, D> cand;
      if constexpr (MT::CURV) {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          cand.w[k] = cw[m * NS + k];
#pragma unroll
          for (int c = 0; c < 3; ++c) cand.e[k][c] = ce[m * NS + k][c];
        }
      } else {
        tti_sub<S, D>(D0, D1, diag[m][0], diag[m][1], cand.w, cand.e);
      }
#pragma unroll
      for (int k = 0; k < NS; ++k) cand.shift[k] = S(0);
      S cv[NS], cw[NS], cv2[NS];
      int cs[NS];
      int64_t cn[NS];
      gather(cand, cv, cw, cs, cn, a.order2 ? cv2 : nullptr);
      S vs[NS], ws[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        // (curvature models: weight-0 offsets left out, as in the update)
        if constexpr (MT::CURV) cv[k] = cw[k] > S(0) ? cv[k] : inf_v<S>();
        vs[k] = cv[k];
        ws[k] = cw[k];
      }
      unsigned cmask = 0;
      if (a.order2) {
        cmask = order2_correct<S, NS>(u_old, a.order2_thr, vs, cv2, ws);
#pragma unroll
        for (int k = 0; k < NS; ++k) cv[k] = vs[k];  // corrected values (unsorted copy)
      }
      const S um = hfm_solve<S, NS>(vs, ws, rhs);
      if (mix_min == (val > um)) {
        val = um;
        st = cand;
        o2mask = cmask;
#pragma unroll
        for (int k = 0; k < NS; ++k) { v[k] = cv[k]; w[k] = cw[k]; side[k] = cs[k]; nb[k] = cn[k]; }
      }
    }
  } else {
    make_stencil<S, MODEL, D>(geo, a.iso_w, st);
    S v2[NS];
    gather(st, v, w, side, nb, a.order2 ? v2 : nullptr);
    S vs[NS], ws[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) { vs[k] = v[k]; ws[k] = w[k]; }
    if (a.order2) {
      o2mask = order2_correct<S, NS>(u_old, a.order2_thr, vs, v2, ws);
#pragma unroll
      for (int k = 0; k < NS; ++k) v[k] = vs[k];  // corrected values (unsorted copy)
    }
    val = hfm_solve<S, NS>(vs, ws, rhs);
  }

  S sum = S(0);
  const bool live = !seed && val != inf_v<S>();
#pragma unroll
  for (int c = 0; c < 3; ++c) out.vec[c] = S(0);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const S diff = val - v[k];
    S f = live ? w[k] * (diff > S(0) ? diff : S(0)) : S(0);
    if ((o2mask >> k) & 1u) f *= S(1.5);
    out.fw[k] = f;
    out.nb[k] = nb[k];
    sum += f;
    const int sg = 2 * side[k] - 1;
#pragma unroll
    for (int c = 0; c < D; ++c) out.vec[c] += f * S(sg * st.e[k][c]);
  }
  out.sum = sum;
  out.u_old = u_old;
  out.val = val;
  out.seed = seed;
}

template <class S, int MODEL, int D, int T0, int T1, int T2>
__global__ void __launch_bounds__(T0* T1* T2) eik_flow(FlowArgs<S> a) {
  using TL = Tile<D, T0, T1, T2>;
  using MT = ModelTraits<MODEL, D>;
  constexpr int NI = TL::NI;
  constexpr int NS = MT::NSYM;
  const int tid = threadIdx.x;
  const int n_o = blockIdx.x;
  int xi[3], xo[3], xt[3] = {0, 0, 0};
  TL::pos(tid, xi);
  tile_pos<D>(n_o, a.g, xo);
#pragma unroll
  for (int k = 0; k < D; ++k) xt[k] = xo[k] * TL::T[k] + xi[k];
  const int64_t n_t = int64_t(n_o) * NI + tid;
  const int64_t N = a.g.size_tot;
  CellFlow<S, NS> r;
  flow_cell<S, MODEL, D, TL>(a, xt, n_t, r);
  const S u_old = r.u_old, val = r.val;
  const bool seed = r.seed;
  const S* fw = r.fw;
  const int64_t* nb = r.nb;
  const S sum = r.sum;
  const S* vec = r.vec;
  if (a.violations) {
    const S tol = a.chk_atol + a.chk_rtol * abs_(u_old);
    const bool bad = !seed && (a.stale ? val < u_old - tol : val > u_old + tol);
    if (bad && a.stale_bits) atomicOr(a.stale_bits + (n_t >> 5), 1u << (n_t & 31));
    const uint64_t m = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(a.violations, int(__popcll(m)));
    if (a.repair) {
      __syncthreads();  // the tile's own reads are done before any store
      if (bad) a.u[n_t] = inf_v<S>();
    }
    return;
  }
  __syncthreads();  // every lane has read its neighbours before any store
  if (!seed && val < u_old) a.u[n_t] = val;
  if (a.fw) {
#pragma unroll
    for (int k = 0; k < NS; ++k) a.fw[k * N + n_t] = fw[k];
  }
  if (a.fidx) {
#pragma unroll
    for (int k = 0; k < NS; ++k)
      a.fidx[k * N + n_t] = (fw[k] != S(0) && nb[k] >= 0) ? int32_t(nb[k]) : INT32_MAX;
  }
  if (a.wsum) a.wsum[n_t] = sum;
  if (a.fvec) {
#pragma unroll
    for (int c = 0; c < D; ++c) a.fvec[c * N + n_t] = vec[c];
  }
}

// ------------------------------------------------------------------ geodesics
// Flow recomputed at each cell the geodesic ODE enters
// (geodesic_online_flow, GeodesicODE_Opt.h:101-238): the flow pass's cell
// computation on the solver's values, geometry and constants, so that no
// flow arrays are kept.
template <class S, int MODEL, int D, class TL> struct GeoFlowOnline {
  const FlowArgs<S>* fa;
  const GeoGrid* g;
  __device__ void at(int64_t n, S (&vec)[D], S& ws) const {
    int y[3] = {0, 0, 0};
    geo_pos<D>(n, *g, y);
    CellFlow<S, ModelTraits<MODEL, D>::NSYM> r;
    flow_cell<S, MODEL, D, TL>(*fa, y, n, r);
    ws = r.sum;
#pragma unroll
    for (int k = 0; k < D; ++k) vec[k] = r.vec[k];
  }
};

template <class S, int MODEL, int D, int T0, int T1, int T2>
__global__ void __launch_bounds__(64) eik_geodesic_online(GeoArgs<S> a, FlowArgs<S> fa) {
  const GeoFlowOnline<S, MODEL, D, Tile<D, T0, T1, T2>> src{&fa, &a.g};
  geodesic_run<S, D>(a, src);
}

// ------------------------------------------------------------------ tables
// Model kernels of one (model, dimension, tile shape); null entries are
// variants that do not exist for the model.
template <class S> struct ModelKernels {
  void (*update)(EikArgs<S>);
  void (*update_walls)(EikArgs<S>);
  void (*update_pre)(EikArgs<S>);          // null: no precomputed-scheme variant
  void (*update_batch)(EikArgs<S>);        // null: no batched variant (FP64)
  void (*update_order2)(EikArgs<S>);       // null: no second-order variant (TTI)
  void (*scheme)(SchemeArgs<S>);
  void (*scheme_general)(SchemeArgs<S>);
  void (*flow)(FlowArgs<S>);
  void (*tti_factor)(TTIFactorArgs<S>);  // TTI only: shift table of the factoring ball
  void (*geodesic_online)(GeoArgs<S>, FlowArgs<S>);  // geodesics with the flow recomputed
  int ni;               // cells per tile
  int block;            // lanes per workgroup of the update kernels
  int tiles_per_block;  // tiles a workgroup visits at once (tiles below a wave)
};

template <class S, int MODEL, int D, int T0, int T1, int T2>
ModelKernels<S> make_model_kernels() {
  ModelKernels<S> k;
  k.ni = T0 * T1 * T2;
  k.block = block_lanes<T0 * T1 * T2>();
  k.tiles_per_block = tiles_per_block<T0 * T1 * T2>();
  if constexpr (T0 * T1 * T2 < 64) {
    // tiles below a wave: the batched FP32 variant only
    k.update = k.update_walls = k.update_pre = k.update_order2 = nullptr;
    k.scheme = k.scheme_general = nullptr;
    k.flow = nullptr;
    k.tti_factor = nullptr;
    k.geodesic_online = nullptr;
    if constexpr (std::is_same<S, float>::value && !ModelTraits<MODEL, D>::STORED)
      k.update_batch = eik_update<S, MODEL, D, T0, T1, T2, VAR_BATCH>;
    else
      k.update_batch = nullptr;
  } else if constexpr (ModelTraits<MODEL, D>::CURV) {
    // Curvature models (3D (x, y, theta) grids): precomputed scheme only,
    // resolved by the general variant (periodic theta axis).
    if constexpr (D == 3) {
      k.update = k.update_walls = k.update_order2 = k.update_batch = nullptr;
      k.update_pre = eik_update<S, MODEL, D, T0, T1, T2, VAR_PRE>;
      k.scheme = k.scheme_general = eik_scheme<S, MODEL, D, T0, T1, T2, true>;
      k.flow = eik_flow<S, MODEL, D, T0, T1, T2>;
      k.geodesic_online = eik_geodesic_online<S, MODEL, D, T0, T1, T2>;
    } else {
      k.update = k.update_walls = k.update_order2 = k.update_batch = k.update_pre = nullptr;
      k.scheme = k.scheme_general = nullptr;
      k.flow = nullptr;
      k.geodesic_online = nullptr;
    }
    k.tti_factor = nullptr;
  } else {
    if constexpr (MODEL == AGD_TTI) {
      // TTI runs from its precomputed scheme only.
      k.update = nullptr;
      k.update_walls = nullptr;
      k.update_batch = nullptr;
      k.update_order2 = nullptr;
    } else {
      k.update = eik_update<S, MODEL, D, T0, T1, T2, VAR_FAST>;
      k.update_walls = eik_update<S, MODEL, D, T0, T1, T2, VAR_GENERAL>;
      k.update_order2 = eik_update<S, MODEL, D, T0, T1, T2, VAR_ORDER2>;
      if constexpr (std::is_same<S, float>::value)
        k.update_batch = eik_update<S, MODEL, D, T0, T1, T2, VAR_BATCH>;
      else
        k.update_batch = nullptr;
    }
    // Precomputed schemes pay off where the stencil is costly to build
    // (Selling decompositions: one for Riemann, seven for TTI, which has no
    // other variant); the drift of Rander is not stored.
    if constexpr (MODEL == AGD_RIEMANN || MODEL == AGD_TTI) {
      k.update_pre = eik_update<S, MODEL, D, T0, T1, T2, VAR_PRE>;
      k.scheme = eik_scheme<S, MODEL, D, T0, T1, T2, false>;
      k.scheme_general = eik_scheme<S, MODEL, D, T0, T1, T2, true>;
    } else {
      k.update_pre = nullptr;
      k.scheme = nullptr;
      k.scheme_general = nullptr;
    }
    k.flow = eik_flow<S, MODEL, D, T0, T1, T2>;
    k.geodesic_online = eik_geodesic_online<S, MODEL, D, T0, T1, T2>;
    if constexpr (MODEL == AGD_TTI) k.tti_factor = eik_tti_factor<S, D, T0, T1, T2>;
    else k.tti_factor = nullptr;
  }
  return k;
}

inline int tile_key(int d, const int32_t* si) {
  return d == 2 ? (si[0] * 100 + si[1]) : (si[0] * 10000 + si[1] * 100 + si[2]);
}

// The compiled tile shapes: 2D 8x8, 16x16, 8x16; 3D 4x4x4, 8x8x4, 4x4x8;
// and, for batched solves only, 2D 4x4, 4x8, 8x4.
template <class S, int MODEL>
bool lookup_model_kernels(int d, const int32_t* si, ModelKernels<S>& out) {
  const int key = tile_key(d, si);
  if (d == 2) {
    switch (key) {
      case 808: out = make_model_kernels<S, MODEL, 2, 8, 8, 1>(); return true;
      case 1616: out = make_model_kernels<S, MODEL, 2, 16, 16, 1>(); return true;
      case 816: out = make_model_kernels<S, MODEL, 2, 8, 16, 1>(); return true;
      case 404: out = make_model_kernels<S, MODEL, 2, 4, 4, 1>(); return true;
      case 408: out = make_model_kernels<S, MODEL, 2, 4, 8, 1>(); return true;
      case 804: out = make_model_kernels<S, MODEL, 2, 8, 4, 1>(); return true;
    }
  } else if (d == 3) {
    switch (key) {
      case 40404: out = make_model_kernels<S, MODEL, 3, 4, 4, 4>(); return true;
      case 80804: out = make_model_kernels<S, MODEL, 3, 8, 8, 4>(); return true;
      case 40408: out = make_model_kernels<S, MODEL, 3, 4, 4, 8>(); return true;
    }
  }
  return false;
}

// Defined (and instantiated for float and double) in eik_<model>.hip.
template <class S> bool lookup_isotropic(int d, const int32_t* si, ModelKernels<S>& out);
template <class S> bool lookup_diagonal(int d, const int32_t* si, ModelKernels<S>& out);
template <class S> bool lookup_riemann(int d, const int32_t* si, ModelKernels<S>& out);
template <class S> bool lookup_rander(int d, const int32_t* si, ModelKernels<S>& out);
template <class S> bool lookup_tti(int d, const int32_t* si, ModelKernels<S>& out);
template <class S> bool lookup_curvature(int model, int d, const int32_t* si, ModelKernels<S>& out);

}  // namespace agd
