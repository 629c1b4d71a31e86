// eikonal.hip -- the eikonal tile solver for gfx950: grid-level worklist
// driver, the fused tile-update kernel, the flow export kernel and the
// seed / wall / stencil helpers.
//
// Tile update (reference: Update.h:20-227, HFMIter.h:10-86,
// Propagation.h:18-167, FiniteDifferences.h:17-106):
//   one workgroup per active tile, one lane per cell.  The lane loads its
//   value, rhs, seed bit and geometry (coalesced: the tiled layout makes a
//   tile contiguous), builds its stencil in registers (Selling for
//   Riemann/Rander), then resolves each of its 2*nsym neighbours once: an LDS
//   slot when the neighbour lies in the tile, otherwise the neighbour's value
//   gathered from HBM/L2 and held in a register for the whole visit (values
//   outside the tile do not change during the in-tile sweeps).  The tile's
//   own values live in LDS and are relaxed niter_i times (Jacobi inside a
//   sweep, early exit once a sweep changes nothing -- which cannot alter the
//   result since the next sweep would read identical inputs).  The change
//   test |u_old-u_new| <= atol + rtol|u_new| is reduced with a block-wide
//   OR (a single wave64 ballot for 64-cell tiles) and drives tile activation.
//
// Worklist (replaces the host round trip of _solvers.py:122-193): the tile
// list of pass k+1 is built on the device while pass k runs.  A tile is
// appended at most once per pass thanks to a per-tile stamp
// (atomicMax(stamp, k+1) < k+1 <=> first to tag it), so neither the list
// nor the flags need clearing.  Three rotating counters: pass k reads
// count[k%3], appends to count[(k+1)%3] and zeroes count[(k+2)%3], which no
// later pass of this launch touches before pass k+1.  Every launch uses the
// same grid; workgroups stride over the list.  The host reads the next
// count every `check_interval` passes (one small copy + sync).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "worklist.hpp"

namespace agd {

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
};

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

// Resolves the 2*nsym neighbours x +- e_k of a cell into one code each:
//   code >= 0 : tiled index of a neighbour outside the tile (gathered),
//   code <= -2: LDS slot (-2 - code) inside the tile,
//   code == -1: no neighbour (outside the domain or hidden by a wall).
template <class S, int MODEL, int D, class TL, bool GENERAL>
__device__ inline void resolve_codes(const Stencil<S, MODEL, D>& st, const int* xi, const int* xt,
                                     const int* xo, int wx, const uint8_t* wall_s,
                                     const uint8_t* wall, const GridDesc& g, int* code,
                                     int tile_base = 0) {
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
        for (int c = 1; c < D; ++c) to = to * g.shape_o[c] + yo[c];
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
      }
    }
  }
}

// Kernel variants: VAR_FAST assumes no walls and no periodic axes, so its
// neighbour indexing carries no wrap-around code; VAR_GENERAL handles both;
// VAR_PRE reads stencil weights and neighbour codes precomputed once per
// solve (any walls / periodicity are already resolved in the codes);
// VAR_BATCH is VAR_FAST over the concatenated tiles of several instances,
// each tile looking up its instance's grid and constants.
enum { VAR_FAST = 0, VAR_GENERAL = 1, VAR_PRE = 2, VAR_BATCH = 3 };

template <class S, int MODEL, int D, int T0, int T1, int T2, int VAR>
__global__ void __launch_bounds__(T0* T1* T2) eik_update(EikArgs<S> a) {
  constexpr bool GENERAL = VAR == VAR_GENERAL;
  constexpr bool WALLS = GENERAL;
  constexpr bool BATCH = VAR == VAR_BATCH;
  static_assert((T0 * T1 * T2) % 64 == 0, "tiles are whole waves (pass_begin scans per wave)");
  using TL = Tile<D, T0, T1, T2>;
  using MT = ModelTraits<MODEL, D>;
  constexpr int NI = TL::NI;
  constexpr int NS = MT::NSYM;
  constexpr int NG = MT::NGEOM;
  constexpr int NJ = 2 * NS;
  __shared__ S u_s[NI];
  __shared__ uint8_t wall_s[WALLS ? NI : 1];

  const int tid = threadIdx.x;
  const PassView pv = pass_begin(a.w);
  const int count = pv.count;
  int xi[3];
  TL::pos(tid, xi);
  const int32_t tag = pv.iter + 1;
  const bool fim = a.solver == AGD_FIM;

  for (int li = blockIdx.x; li < count; li += gridDim.x) {
    const int pk = pass_entry(pv, li);
    int xo[3];
    // Grid and constants of the tile's instance (the solve's own when not
    // batched).  Batched list entries are tile ids, else packed coordinates.
    GridDesc gi;
    int tbase = 0;
    S atol, rtol, iso_w[3];
    int n_o;
    if constexpr (BATCH) {
      n_o = __builtin_amdgcn_readfirstlane(pk);
      const int ii = __builtin_amdgcn_readfirstlane(a.tile_inst[n_o]);
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
    const int n_t = n_o * NI + tid;

    // ---- activation inputs first: the tile and its face neighbours
    // (lanes 0..2D of wave 0), and their FIM scores of the previous pass.
    int nb = -1, nbp = pk, sc = 0;
    if (tid < 2 * D + 1) {
      if (tid == 2 * D) {
        nb = n_o;
      } else {
        nb = tile_neighbor<D>(xo, tid >> 1, 2 * (tid & 1) - 1, gi, &nbp);
        if (BATCH && nb >= 0) {
          nb += tbase;
          nbp = nb;
        }
      }
      if (fim && nb >= 0) sc = read_score(pv.score_prev, nb, pv.iter - 1);
    }

    // ---- cell loads (coalesced: a tile is contiguous)
    const S u_old = a.u[n_t];
    const S rhs = a.rhs[n_t];
    const bool seed = (a.seeds[n_t >> 5] >> (n_t & 31)) & 1u;
    int wx = 255;
    if constexpr (WALLS) {
      if (a.wall) wx = a.wall[n_t];
      wall_s[tid] = uint8_t(wx);
    }
    u_s[tid] = seed ? rhs : u_old;

    int code[NJ];
    S sw[NS], shift[NS];
    if constexpr (MT::NMIX > 1) {
      // several sub-schemes: read per sweep below
    } else if constexpr (VAR == VAR_PRE) {
      // Stencil weights and neighbour codes precomputed for the whole solve.
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        sw[k] = a.sw[k * a.g.size_tot + n_t];
        shift[k] = S(0);
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) code[j] = a.scode[j * a.g.size_tot + n_t];
    } else {
      S geo[NG > 0 ? NG : 1];
#pragma unroll
      for (int k = 0; k < NG; ++k) {
        const S* __restrict__ gk = a.geom + k * a.g.size_tot;  // uniform base
        geo[k] = gk[n_t];
      }
      Stencil<S, MODEL, D> st;
      make_stencil<S, MODEL, D>(geo, iso_w, st);
      if constexpr (WALLS) __syncthreads();  // wall_s complete before the visibility walks
      resolve_codes<S, MODEL, D, TL, GENERAL>(st, xi, xt, xo, wx, wall_s, a.wall, gi, code, tbase);
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        sw[k] = st.w[k];
        shift[k] = st.shift[k];
      }
    }
    if constexpr (MT::NMIX > 1) {
      // ---- TTI: the min or max over the sub-schemes of their upwind
      // updates (HFMIter.h:29-56, strict in-tile iteration).  Weights and
      // neighbour codes are re-read at every sweep (cache hits) rather than
      // held in registers; out-of-tile values are read at every sweep.
      const int64_t N = a.g.size_tot;
      const bool mix_min = a.scode[int64_t(MT::NMIX * NJ) * N + n_t] != 0;
      __syncthreads();  // u_s complete
      const bool active = !seed;
      for (int it = 0; it < a.niter_i; ++it) {
        S best = mix_min ? inf_v<S>() : -inf_v<S>();
#pragma unroll 1
        for (int m = 0; m < MT::NMIX; ++m) {
          S v[NS], w[NS];
#pragma unroll
          for (int k = 0; k < NS; ++k) {
            w[k] = a.sw[int64_t(m * NS + k) * N + n_t];
            S side[2];
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
              const int c = a.scode[int64_t(m * NJ + 2 * k + sd) * N + n_t];
              side[sd] = c <= -2 ? u_s[-2 - c] : (c >= 0 ? a.u[c] : inf_v<S>());
            }
            v[k] = (side[0] > side[1]) ? side[1] : side[0];
          }
          const S um = hfm_solve<S, NS>(v, w, rhs);
          if (mix_min == (best > um)) best = um;
        }
        __syncthreads();
        const bool ch = active && (u_s[tid] > best);
        if (ch) u_s[tid] = best;
        if (!block_any<NI>(ch)) break;
      }
    } else {
      // All gathers are issued before any is consumed (safe address = own
      // cell), so a visit costs a single memory round trip.
      S g[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) g[j] = a.u[code[j] >= 0 ? code[j] : n_t];
      S vo[NJ];
#pragma unroll
      for (int k = 0; k < NS; ++k) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int j = 2 * k + s;
          const S sh = (s == 0) ? shift[k] : -shift[k];
          vo[j] = code[j] <= -2 ? sh : (code[j] >= 0 ? g[j] + sh : inf_v<S>());
        }
      }
      __syncthreads();  // u_s complete

      // ---- in-tile sweeps (HFMIter.h:25-84, DECREASING)
      const bool active = !seed;
      for (int it = 0; it < a.niter_i; ++it) {
        S v[NS], w[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          const S a0 = code[2 * k] <= -2 ? u_s[-2 - code[2 * k]] + vo[2 * k] : vo[2 * k];
          const S a1 = code[2 * k + 1] <= -2 ? u_s[-2 - code[2 * k + 1]] + vo[2 * k + 1] : vo[2 * k + 1];
          v[k] = (a0 > a1) ? a1 : a0;  // Greater(v,-) keeps the -e side on ties
          w[k] = sw[k];
        }
        const S un = hfm_solve<S, NS>(v, w, rhs);
        __syncthreads();
        const bool ch = active && (u_s[tid] > un);
        if (ch) u_s[tid] = un;
        if (!block_any<NI>(ch)) break;
      }
    }

    // ---- store and convergence (Update.h:163-219)
    const S un = u_s[tid];
    a.u[n_t] = un;
    const S diff = abs_(u_old - un);
    const S tol = atol + rtol * abs_(un);
    const bool changed = !(isnan_(diff) || diff <= tol);
    const bool any = block_any<NI>(changed);

    // ---- activation (Propagation.h:103-167), wave 0 only.  The
    // global-iteration solver runs with AGSI tags: they only decide when to
    // stop, every pass visits the full list.
    if (tid < 64) {
      if (!fim) {
        enqueue_wave(any && nb >= 0, nb, nbp, tag, a.w.stamp, pv.list_out, pv.count_out);
      } else {
        const int front = a.fim_front;
        const int prev = __shfl(sc, 2 * D);
        int mx = 0;
#pragma unroll
        for (int i = 0; i < 2 * D; ++i) mx = max(mx, __shfl(sc, i));
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
        if (tid == 2 * D) pv.score_next[n_o] = (pv.iter << 8) | next;
        const bool want = (tid == 2 * D) ? any : (prop && tid < 2 * D && nb >= 0);
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
  if constexpr (MT::NMIX > 1) {
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
      resolve_codes<S, MODEL, D, TL, GENERAL>(st, xi, xt, xo, wx, wall_s, a.wall, a.g, code);
#pragma unroll
      for (int k = 0; k < NS; ++k) a.sw[int64_t(m * NS + k) * N + n_t] = st.w[k];
#pragma unroll
      for (int j = 0; j < 2 * NS; ++j) a.scode[int64_t(m * 2 * NS + j) * N + n_t] = code[j];
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

// Initial list: tiles holding at least one trigger bit (or every tile).
// 64-lane workgroups, one tile per lane, wave-aggregated appends into
// sub-list blockIdx % kShards of pass 0.
template <int D, int NI>
__global__ void __launch_bounds__(64) eik_init_list(const uint32_t* trigger, GridDesc g, int32_t all,
                                                   int32_t* stamp, int32_t* lists, int32_t* counts,
                                                   int32_t shard_cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool on = false;
  int xo[3] = {0, 0, 0};
  if (t < g.ntiles) {
    on = all != 0;
    const int64_t b0 = int64_t(t) * NI, b1 = b0 + NI;
    for (int64_t b = b0; b < b1 && !on; ) {
      const int64_t wi = b >> 5;
      const int lo = int(b & 31);
      const int64_t lim = std::min<int64_t>(b1, (wi + 1) << 5);
      const int hi = int(lim - (wi << 5));
      uint32_t word = trigger[wi];
      uint32_t mask = (hi >= 32 ? 0xffffffffu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
      on = (word & mask) != 0;
      b = lim;
    }
    tile_pos<D>(t, g, xo);
  }
  const int shard = blockIdx.x % kShards;
  enqueue_wave(on, on ? t : 0, pack_tile<D>(xo), 0, stamp, lists + int64_t(shard) * shard_cap,
               counts + shard * kCountStride);
}

// Batched solves: the same with tile ids as entries.
template <int NI>
__global__ void __launch_bounds__(64) init_list_ids(const uint32_t* trigger, int32_t ntiles, int32_t all,
                                                   int32_t* stamp, int32_t* lists, int32_t* counts,
                                                   int32_t shard_cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool on = false;
  if (t < ntiles) {
    on = all != 0;
    const int64_t b0 = int64_t(t) * NI, b1 = b0 + NI;
    for (int64_t b = b0; b < b1 && !on; b += 32) {
      // NI is a multiple of 64: whole words
      on = trigger[b >> 5] != 0u;
    }
  }
  const int shard = blockIdx.x % kShards;
  enqueue_wave(on, on ? t : 0, t, 0, stamp, lists + int64_t(shard) * shard_cap,
               counts + shard * kCountStride);
}

// Every tile id (batched global-iteration passes).
__global__ void all_ids_kernel(int32_t ntiles, int32_t* list, int32_t* count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) list[t] = t;
  if (t == 0) *count = ntiles;
}

// tile -> instance map: the last instance whose tile_base is <= t.
__global__ void tile_inst_kernel(int32_t ninst, const InstDesc* inst, int32_t ntiles,
                                 int32_t* tile_inst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (inst[mid].tile_base <= t) lo = mid;
    else hi = mid - 1;
  }
  tile_inst[t] = lo;
}

// Every tile, as packed coordinates (global-iteration passes).
template <int D>
__global__ void all_tiles_kernel(GridDesc g, int32_t* list, int32_t* count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < g.ntiles) {
    int xo[3];
    tile_pos<D>(t, g, xo);
    list[t] = pack_tile<D>(xo);
  }
  if (t == 0) *count = g.ntiles;
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
};

template <class S, int MODEL, int D, int T0, int T1, int T2>
__global__ void __launch_bounds__(T0* T1* T2) eik_flow(FlowArgs<S> a) {
  using TL = Tile<D, T0, T1, T2>;
  using MT = ModelTraits<MODEL, D>;
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
  const int64_t n_t = int64_t(n_o) * NI + tid;
  const int64_t N = a.g.size_tot;

  const S u_old = a.u[n_t];
  const S rhs = a.rhs[n_t];
  const bool seed = (a.seeds[n_t >> 5] >> (n_t & 31)) & 1u;
  S geo[NG > 0 ? NG : 1];
#pragma unroll
  for (int k = 0; k < NG; ++k) geo[k] = a.geom[k * N + n_t];
  const int wx = a.wall ? int(a.wall[n_t]) : 255;

  // Neighbour values (min over -e/+e, ties keep -e), sides and indices of a
  // stencil.
  auto gather = [&](const Stencil<S, MODEL, D>& st, S* v, S* w, int* side, int64_t* nb) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      S best = inf_v<S>();
      int bs = 0;
      int64_t bi = -1;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int sg = 2 * s - 1;
        int e[3] = {sg * st.e[k][0], sg * st.e[k][1], sg * st.e[k][2]};
        const S sh = (s == 0) ? st.shift[k] : -st.shift[k];
        bool vis = true;
        if (a.wall && wx != 255) {
          vis = visible<D>(e, wx, [&](const int* ww) -> int {
            int yt[3] = {xt[0] + ww[0], xt[1] + ww[1], xt[2] + ww[2]};
            const int64_t m = index_tot<D, TL>(yt, a.g);
            return m >= 0 ? int(a.wall[m]) : 255;
          });
        }
        int yt[3] = {xt[0] + e[0], xt[1] + e[1], xt[2] + e[2]};
        const int64_t m = vis ? index_tot<D, TL>(yt, a.g) : -1;
        const S val = (m >= 0) ? a.u[m] + sh : inf_v<S>();
        if (s == 0) { best = val; bs = 0; bi = m; }
        else if (best > val) { best = val; bs = 1; bi = m; }
      }
      v[k] = best;
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
  if constexpr (MT::NMIX > 1) {
    // TTI: keep the sub-scheme whose update wins the min/max (HFMIter.h:48-54).
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
    S diag[kTTIMix][2];
    const bool mix_min = tti_diags<S>(geo, geo + 2, diag);
    S D0[NS], D1[NS];
    tti_base<S, D>(geo, D0, D1);
    val = mix_min ? inf_v<S>() : -inf_v<S>();
#pragma unroll 1
    for (int m = 0; m < MT::NMIX; ++m) {
      Stencil<S, MODEL, D> cand;
      tti_sub<S, D>(D0, D1, diag[m][0], diag[m][1], cand.w, cand.e);
#pragma unroll
      for (int k = 0; k < NS; ++k) cand.shift[k] = S(0);
      S cv[NS], cw[NS];
      int cs[NS];
      int64_t cn[NS];
      gather(cand, cv, cw, cs, cn);
      S vs[NS], ws[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) { vs[k] = cv[k]; ws[k] = cw[k]; }
      const S um = hfm_solve<S, NS>(vs, ws, rhs);
      if (mix_min == (val > um)) {
        val = um;
        st = cand;
#pragma unroll
        for (int k = 0; k < NS; ++k) { v[k] = cv[k]; w[k] = cw[k]; side[k] = cs[k]; nb[k] = cn[k]; }
      }
    }
  } else {
    make_stencil<S, MODEL, D>(geo, a.iso_w, st);
    gather(st, v, w, side, nb);
    S vs[NS], ws[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) { vs[k] = v[k]; ws[k] = w[k]; }
    val = hfm_solve<S, NS>(vs, ws, rhs);
  }

  S fw[NS];
  S sum = S(0);
  S vec[3] = {S(0), S(0), S(0)};
  const bool live = !seed && val != inf_v<S>();
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const S diff = val - v[k];
    fw[k] = live ? w[k] * (diff > S(0) ? diff : S(0)) : S(0);
    sum += fw[k];
    const int sg = 2 * side[k] - 1;
#pragma unroll
    for (int c = 0; c < D; ++c) vec[c] += fw[k] * S(sg * st.e[k][c]);
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

// ------------------------------------------------------------------ stencil export
template <class S, int MODEL, int D>
__global__ void stencil_export_kernel(int64_t n, const S* geom, S iw0, S iw1, S iw2, S* wout,
                                      int32_t* eout) {
  using MT = ModelTraits<MODEL, D>;
  constexpr int NS = MT::NSYM;
  constexpr int NG = MT::NGEOM;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  S geo[NG > 0 ? NG : 1];
#pragma unroll
  for (int k = 0; k < NG; ++k) geo[k] = geom[k * n + c];
  S iw[3] = {iw0, iw1, iw2};
  Stencil<S, MODEL, D> st;
  make_stencil<S, MODEL, D>(geo, iw, st);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    wout[c * NS + k] = st.w[k];
#pragma unroll
    for (int l = 0; l < D; ++l) eout[(c * NS + k) * D + l] = st.e[k][l];
  }
}

// ------------------------------------------------------------------ seeds / walls
template <class S>
__global__ void scatter_seeds_kernel(int64_t n, const int64_t* idx, const double* val, S* rhs,
                                     uint32_t* seeds, uint32_t* trigger) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t m = idx[i];
  rhs[m] = S(val[i]);
  atomicOr(seeds + (m >> 5), 1u << (m & 31));
  if (trigger) atomicOr(trigger + (m >> 5), 1u << (m & 31));
}

template <class S>
__global__ void walls_mark_kernel(int64_t N, const uint8_t* walls, S* rhs, uint32_t* seeds,
                                  uint8_t* dist) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool w = walls[i] != 0;
  dist[i] = w ? 0 : 255;
  if (w) {
    rhs[i] = inf_v<S>();
    atomicOr(seeds + (i >> 5), 1u << (i & 31));
  }
}

// One min-plus pass with the L1 unit kernel (Kernel_InfConvolution.h:66-91
// with distance_kernel(1, ndim, ord=1)), saturated at 255.  Operates on the
// tiled layout; padded cells stay at 255.
template <int D, int T0, int T1, int T2>
__global__ void wall_dist_pass(GridDesc g, const uint8_t* in, uint8_t* out) {
  using TL = Tile<D, T0, T1, T2>;
  const int n_o = blockIdx.x;
  const int tid = threadIdx.x;
  int xi[3], xo[3], xt[3] = {0, 0, 0};
  TL::pos(tid, xi);
  tile_pos<D>(n_o, g, xo);
  bool dom = true;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xt[k] = xo[k] * TL::T[k] + xi[k];
    dom = dom && xt[k] < g.shape_tot[k];
  }
  const int64_t n_t = int64_t(n_o) * TL::NI + tid;
  if (!dom) { out[n_t] = 255; return; }
  int best = in[n_t];
  // 3^D neighbourhood, L1 kernel weights.
  const int r2 = (D == 3) ? 1 : 0;
  for (int d0 = -1; d0 <= 1; ++d0)
    for (int d1 = -1; d1 <= 1; ++d1)
      for (int d2 = -r2; d2 <= r2; ++d2) {
        int y[3] = {xt[0] + d0, xt[1] + d1, xt[2] + d2};
        const int64_t m = index_tot<D, TL>(y, g);
        if (m < 0) continue;
        const int kw = (d0 < 0 ? -d0 : d0) + (d1 < 0 ? -d1 : d1) + (d2 < 0 ? -d2 : d2);
        const int vv = in[m];
        const int p = (vv >= 255 - kw) ? 255 : vv + kw;
        best = p < best ? p : best;
      }
  out[n_t] = uint8_t(best);
}

__global__ void wall_dist_cap(int64_t N, uint8_t* d, int bound) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N && d[i] > bound) d[i] = 255;
}

// ------------------------------------------------------------------ dispatch
template <class S> struct KernelSet {
  void (*update)(EikArgs<S>);
  void (*update_walls)(EikArgs<S>);
  void (*update_pre)(EikArgs<S>);          // null: no precomputed-scheme variant
  void (*update_batch)(EikArgs<S>);        // null: no batched variant (FP64)
  void (*init_ids)(const uint32_t*, int32_t, int32_t, int32_t*, int32_t*, int32_t*, int32_t);
  void (*scheme)(SchemeArgs<S>);
  void (*scheme_general)(SchemeArgs<S>);
  void (*flow)(FlowArgs<S>);
  void (*init)(const uint32_t*, GridDesc, int32_t, int32_t*, int32_t*, int32_t*, int32_t);
  void (*all_tiles)(GridDesc, int32_t*, int32_t*);
  void (*wallpass)(GridDesc, const uint8_t*, uint8_t*);
  int ni;
};

template <class S, int MODEL, int D, int T0, int T1, int T2>
KernelSet<S> make_set() {
  KernelSet<S> k;
  if constexpr (MODEL == AGD_TTI) {
    // TTI runs from its precomputed scheme only.
    k.update = nullptr;
    k.update_walls = nullptr;
    k.update_batch = nullptr;
  } else {
    k.update = eik_update<S, MODEL, D, T0, T1, T2, VAR_FAST>;
    k.update_walls = eik_update<S, MODEL, D, T0, T1, T2, VAR_GENERAL>;
    if constexpr (std::is_same<S, float>::value)
      k.update_batch = eik_update<S, MODEL, D, T0, T1, T2, VAR_BATCH>;
    else
      k.update_batch = nullptr;
  }
  k.init_ids = init_list_ids<T0 * T1 * T2>;
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
  k.init = eik_init_list<D, T0 * T1 * T2>;
  k.all_tiles = all_tiles_kernel<D>;
  k.wallpass = wall_dist_pass<D, T0, T1, T2>;
  k.ni = T0 * T1 * T2;
  return k;
}

static int tile_key(int d, const int32_t* si) {
  return d == 2 ? (si[0] * 100 + si[1]) : (si[0] * 10000 + si[1] * 100 + si[2]);
}

template <class S, int MODEL>
static bool lookup_model(int d, const int32_t* si, KernelSet<S>& out) {
  const int key = tile_key(d, si);
  if (d == 2) {
    switch (key) {
      case 808: out = make_set<S, MODEL, 2, 8, 8, 1>(); return true;
      case 1616: out = make_set<S, MODEL, 2, 16, 16, 1>(); return true;
      case 816: out = make_set<S, MODEL, 2, 8, 16, 1>(); return true;
    }
  } else if (d == 3) {
    switch (key) {
      case 40404: out = make_set<S, MODEL, 3, 4, 4, 4>(); return true;
      case 80804: out = make_set<S, MODEL, 3, 8, 8, 4>(); return true;
      case 40408: out = make_set<S, MODEL, 3, 4, 4, 8>(); return true;
    }
  }
  return false;
}

template <class S>
static bool lookup(int model, int d, const int32_t* si, KernelSet<S>& out) {
  switch (model) {
    case AGD_ISOTROPIC: return lookup_model<S, AGD_ISOTROPIC>(d, si, out);
    case AGD_DIAGONAL: return lookup_model<S, AGD_DIAGONAL>(d, si, out);
    case AGD_RIEMANN: return lookup_model<S, AGD_RIEMANN>(d, si, out);
    case AGD_RANDER: return lookup_model<S, AGD_RANDER>(d, si, out);
    case AGD_TTI: return lookup_model<S, AGD_TTI>(d, si, out);
  }
  return false;
}

static int check_desc(const agd_eikonal_desc* d) {
  AGD_REQUIRE(d != nullptr, "null descriptor");
  AGD_REQUIRE(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
  AGD_REQUIRE(d->model >= AGD_ISOTROPIC && d->model <= AGD_TTI, "unknown model");
  AGD_REQUIRE(d->scalar == AGD_F32 || d->scalar == AGD_F64, "unknown scalar type");
  for (int k = 0; k < d->ndim; ++k) {
    AGD_REQUIRE(d->shape[k] >= 1 && d->shape[k] < (1 << 30), "shape out of range");
    AGD_REQUIRE(d->shape_i[k] >= 1, "tile shape must be positive");
  }
  AGD_REQUIRE(d->niter_i >= 1, "niter_i must be >= 1");
  AGD_REQUIRE(d->nitermax_o >= 1, "nitermax_o must be >= 1");
  int64_t size_tot = 1;
  for (int k = 0; k < d->ndim; ++k) size_tot *= ceil_div(d->shape[k], d->shape_i[k]) * d->shape_i[k];
  AGD_REQUIRE(size_tot < (int64_t(1) << 31), "grids of 2^31 padded cells or more are not supported");
  for (int k = 0; k < d->ndim; ++k)
    AGD_REQUIRE(ceil_div(d->shape[k], d->shape_i[k]) <= (d->ndim == 2 ? 32768 : 1024),
                "too many tiles along an axis (2D: 32768, 3D: 1024)");
  return AGD_OK;
}

static GridDesc make_grid(const agd_eikonal_desc* d) {
  GridDesc g{};
  g.ntiles = 1;
  int64_t ni = 1;
  for (int k = 0; k < 3; ++k) {
    g.shape_tot[k] = k < d->ndim ? int32_t(d->shape[k]) : 1;
    g.shape_o[k] = k < d->ndim ? int32_t(ceil_div(d->shape[k], d->shape_i[k])) : 1;
    if (k < d->ndim) {
      g.ntiles *= g.shape_o[k];
      ni *= d->shape_i[k];
    }
  }
  g.periodic = 0;
  for (int k = 0; k < d->ndim; ++k)
    if (d->periodic[k]) g.periodic |= (1 << k);
  g.size_tot = int64_t(g.ntiles) * ni;
  return g;
}

// Solver workspace, cached per device and grown on demand.
static std::mutex g_ws_mutex;
static std::unordered_map<int, Workspace> g_ws;

int get_workspace(int64_t ntiles, int64_t shard_cap, int64_t nhist, Workspace*& out) {
  int dev = 0;
  AGD_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  Workspace& w = g_ws[dev];
  if (!w.ev0) {
    AGD_HIP_TRY(hipEventCreate(&w.ev0));
    AGD_HIP_TRY(hipEventCreate(&w.ev1));
    AGD_HIP_TRY(hipHostMalloc((void**)&w.host_counts, kRow * sizeof(int32_t), hipHostMallocDefault));
    AGD_HIP_TRY(hipStreamCreateWithFlags(&w.capture_stream, hipStreamNonBlocking));
    AGD_HIP_TRY(hipMalloc((void**)&w.counts, kCountsLen * sizeof(int32_t)));
  }
  if (w.tile_cap < ntiles) {
    if (w.stamp) {
      AGD_HIP_TRY(hipFree(w.stamp));
      AGD_HIP_TRY(hipFree(w.score));
    }
    const int64_t cap = std::max<int64_t>(ntiles, 1024);
    AGD_HIP_TRY(hipMalloc((void**)&w.stamp, cap * sizeof(int32_t)));
    AGD_HIP_TRY(hipMalloc((void**)&w.score, 2 * cap * sizeof(int32_t)));
    w.tile_cap = cap;
  }
  if (w.shard_cap < shard_cap) {
    if (w.lists) AGD_HIP_TRY(hipFree(w.lists));
    const int64_t cap = std::max<int64_t>(shard_cap, 256);
    AGD_HIP_TRY(hipMalloc((void**)&w.lists, 2 * kShards * cap * sizeof(int32_t)));
    w.shard_cap = cap;
  }
  if (w.hist_cap < nhist) {
    if (w.hist) AGD_HIP_TRY(hipFree(w.hist));
    const int64_t cap = std::max<int64_t>(nhist, 4096);
    AGD_HIP_TRY(hipMalloc((void**)&w.hist, cap * sizeof(int32_t)));
    w.hist_cap = cap;
  }
  out = &w;
  return AGD_OK;
}

int blocks_for(int ntiles, int lanes_per_block) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t p;
    cus = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
              ? p.multiProcessorCount
              : 256;
  }
  // Enough resident waves per CU to hide the neighbour-gather latency; the
  // workgroups stride over the tile list.
  const int waves_per_block = (lanes_per_block + 63) / 64;
  const int target = cus * std::max(1, 32 / waves_per_block);
  return std::max(1, std::min(ntiles, target));
}

template <class S>
static int solve_impl(const agd_eikonal_desc* d, const agd_eikonal_io* io,
                      agd_eikonal_stats* stats, hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(io && io->values && io->rhs && io->seeds, "values, rhs and seeds are required");
  AGD_REQUIRE(d->model == AGD_ISOTROPIC || io->geom, "geometry required for this model");
  AGD_REQUIRE(d->solver != AGD_FIM || d->fim_front_width >= 2, "fim_front_width must be >= 2");
  const GridDesc g = make_grid(d);
  const int ni = ks.ni;
  const int blocks = blocks_for(g.ntiles, ni);
  const bool global = d->solver == AGD_GLOBAL;

  const bool pre = io->scheme_weights != nullptr || io->scheme_codes != nullptr;
  AGD_REQUIRE(!pre || (io->scheme_weights && io->scheme_codes),
              "scheme_weights and scheme_codes go together");
  AGD_REQUIRE(!pre || ks.update_pre, "no precomputed-scheme variant for this model");
  AGD_REQUIRE(pre || ks.update, "this model runs from a precomputed scheme (agd_eikonal_scheme)");
  EikArgs<S> a{};
  a.u = static_cast<S*>(io->values);
  a.geom = static_cast<const S*>(io->geom);
  a.sw = static_cast<const S*>(io->scheme_weights);
  a.scode = io->scheme_codes;
  a.rhs = static_cast<const S*>(io->rhs);
  a.seeds = io->seeds;
  a.wall = io->wall_dist;
  a.g = g;
  a.atol = S(d->atol);
  a.rtol = S(d->rtol);
  for (int k = 0; k < 3; ++k) a.iso_w[k] = S(d->iso_weights[k]);
  a.niter_i = d->niter_i;
  a.solver = global ? AGD_AGSI : d->solver;
  a.fim_front = d->fim_front_width;

  auto init = [&](Workspace* ws) -> int {
    a.w.lists = ws->lists;
    a.w.counts = ws->counts;
    a.w.stamp = ws->stamp;
    a.w.hist = ws->hist;
    a.w.score = ws->score;
    a.w.tile_cap = int32_t(ws->tile_cap);
    a.w.shard_cap = int32_t(ws->shard_cap);
    a.w.fixed = nullptr;
    AGD_HIP_TRY(hipMemsetAsync(ws->score, 0x7f, size_t(2 * ws->tile_cap) * sizeof(int32_t), stream));
    if (global) {
      // Every pass visits the full list (kept in the score buffer, unused
      // by this solver); the AGSI tags only decide when to stop.
      hipLaunchKernelGGL(ks.all_tiles, dim3(unsigned(ceil_div(g.ntiles, 256))), dim3(256), 0,
                         stream, g, ws->score, ws->counts + kFixedIdx);
      a.w.fixed = ws->score;
    } else {
      const uint32_t* trig = io->trigger ? io->trigger : io->seeds;
      hipLaunchKernelGGL(ks.init, dim3(unsigned(ceil_div(g.ntiles, 64))), dim3(64), 0, stream, trig,
                         g, d->activate_all, ws->stamp, ws->lists, ws->counts, a.w.shard_cap);
    }
    return AGD_OK;
  };
  auto launch = [&](int32_t koff, hipStream_t s) {
    a.w.koff = koff;
    auto kern = pre ? ks.update_pre : ((io->wall_dist || g.periodic) ? ks.update_walls : ks.update);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(ni), 0, s, a);
  };
  WorklistResult r;
  int st = run_worklist(g.ntiles, blocks, d->nitermax_o, d->check_interval, stream, init, launch,
                        r, d->time_launches != 0, d->use_graph != 0);
  if (st) return st;
  if (stats) {
    stats->kernel_ms = r.kernel_ms;
    stats->niter_o = r.niter;
    stats->tile_updates = r.visits;
    stats->cell_updates = r.visits * ni;
    stats->launches = r.passes;
    stats->gpu_ms = r.ms;
    stats->converged = r.converged ? 1 : 0;
    stats->grid_blocks = blocks;
  }
  if (!r.converged) {
    set_error("eikonal solver did not converge within nitermax_o passes");
    return AGD_NOCONV;
  }
  return AGD_OK;
}

template <class S>
static int flow_impl(const agd_eikonal_desc* d, const agd_eikonal_io* io, const agd_flow_io* f,
                     hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(io && io->values && io->rhs && io->seeds && f, "missing buffers");
  const GridDesc g = make_grid(d);
  FlowArgs<S> a{};
  a.u = static_cast<S*>(io->values);
  a.geom = static_cast<const S*>(io->geom);
  a.rhs = static_cast<const S*>(io->rhs);
  a.seeds = io->seeds;
  a.wall = io->wall_dist;
  a.g = g;
  for (int k = 0; k < 3; ++k) a.iso_w[k] = S(d->iso_weights[k]);
  a.fw = static_cast<S*>(f->flow_weights);
  a.wsum = static_cast<S*>(f->flow_weightsum);
  a.fidx = f->flow_indices;
  a.fvec = static_cast<S*>(f->flow_vector);
  hipLaunchKernelGGL(ks.flow, dim3(g.ntiles), dim3(ks.ni), 0, stream, a);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

// Per-device cache of the instance table and tile -> instance map.
struct BatchBuffers {
  InstDesc* inst = nullptr;
  int32_t* tile_inst = nullptr;
  int64_t inst_cap = 0, tile_cap = 0;
};
static std::mutex g_batch_mutex;
static std::unordered_map<int, BatchBuffers> g_batch;

template <class S>
static int solve_batched_impl(const agd_eikonal_desc* d, int32_t ninst,
                              const agd_batch_instance* bi, const agd_eikonal_io* io,
                              agd_eikonal_stats* stats, hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(ks.update_batch, "batched solves are FP32 only");
  AGD_REQUIRE(ninst >= 1 && bi, "at least one instance is required");
  AGD_REQUIRE(io && io->values && io->rhs && io->seeds, "values, rhs and seeds are required");
  AGD_REQUIRE(d->model == AGD_ISOTROPIC || io->geom, "geometry required for this model");
  AGD_REQUIRE(!io->wall_dist && !io->scheme_weights && !io->scheme_codes,
              "batched solves take no walls and no precomputed scheme");
  AGD_REQUIRE(d->solver != AGD_FIM || d->fim_front_width >= 2, "fim_front_width must be >= 2");
  for (int k = 0; k < d->ndim; ++k) AGD_REQUIRE(!d->periodic[k], "batched solves are not periodic");
  const int ni = ks.ni;
  std::vector<InstDesc> h(static_cast<size_t>(ninst));
  int64_t tiles = 0;
  for (int i = 0; i < ninst; ++i) {
    InstDesc& I = h[size_t(i)];
    I = InstDesc{};
    I.g.ntiles = 1;
    for (int k = 0; k < 3; ++k) {
      const int64_t n = k < d->ndim ? bi[i].shape[k] : 1;
      AGD_REQUIRE(n >= 1 && n < (1 << 30), "instance shape out of range");
      const int64_t no = ceil_div(n, k < d->ndim ? d->shape_i[k] : 1);
      AGD_REQUIRE(no <= (d->ndim == 2 ? 32768 : 1024), "too many tiles along an axis");
      I.g.shape_tot[k] = int32_t(n);
      I.g.shape_o[k] = int32_t(no);
      I.g.ntiles *= int32_t(no);
      I.iso_w[k] = bi[i].iso_weights[k];
    }
    I.g.periodic = 0;
    I.g.size_tot = int64_t(I.g.ntiles) * ni;
    I.tile_base = int32_t(tiles);
    I.atol = bi[i].atol;
    I.rtol = bi[i].rtol;
    tiles += I.g.ntiles;
    AGD_REQUIRE(tiles * ni < (int64_t(1) << 31), "batches of 2^31 padded cells or more are not supported");
  }
  GridDesc gb{};
  gb.ntiles = int32_t(tiles);
  gb.size_tot = tiles * ni;
  for (int k = 0; k < 3; ++k) gb.shape_tot[k] = gb.shape_o[k] = 1;

  BatchBuffers* bb = nullptr;
  {
    int dev = 0;
    AGD_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    bb = &g_batch[dev];
    if (bb->inst_cap < ninst) {
      if (bb->inst) AGD_HIP_TRY(hipFree(bb->inst));
      AGD_HIP_TRY(hipMalloc((void**)&bb->inst, size_t(ninst) * sizeof(InstDesc)));
      bb->inst_cap = ninst;
    }
    if (bb->tile_cap < tiles) {
      if (bb->tile_inst) AGD_HIP_TRY(hipFree(bb->tile_inst));
      AGD_HIP_TRY(hipMalloc((void**)&bb->tile_inst, size_t(tiles) * sizeof(int32_t)));
      bb->tile_cap = tiles;
    }
  }
  AGD_HIP_TRY(hipMemcpyAsync(bb->inst, h.data(), size_t(ninst) * sizeof(InstDesc),
                             hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(tile_inst_kernel, dim3(unsigned(ceil_div(tiles, 256))), dim3(256), 0, stream,
                     ninst, bb->inst, int32_t(tiles), bb->tile_inst);
  AGD_HIP_TRY(hipGetLastError());
  AGD_HIP_TRY(hipStreamSynchronize(stream));  // h is released on return

  const int blocks = blocks_for(int(tiles), ni);
  const bool global = d->solver == AGD_GLOBAL;
  EikArgs<S> a{};
  a.u = static_cast<S*>(io->values);
  a.geom = static_cast<const S*>(io->geom);
  a.rhs = static_cast<const S*>(io->rhs);
  a.seeds = io->seeds;
  a.g = gb;
  a.niter_i = d->niter_i;
  a.solver = global ? AGD_AGSI : d->solver;
  a.fim_front = d->fim_front_width;
  a.inst = bb->inst;
  a.tile_inst = bb->tile_inst;
  auto init = [&](Workspace* ws) -> int {
    a.w.lists = ws->lists;
    a.w.counts = ws->counts;
    a.w.stamp = ws->stamp;
    a.w.hist = ws->hist;
    a.w.score = ws->score;
    a.w.tile_cap = int32_t(ws->tile_cap);
    a.w.shard_cap = int32_t(ws->shard_cap);
    a.w.fixed = nullptr;
    AGD_HIP_TRY(hipMemsetAsync(ws->score, 0x7f, size_t(2 * ws->tile_cap) * sizeof(int32_t), stream));
    if (global) {
      hipLaunchKernelGGL(all_ids_kernel, dim3(unsigned(ceil_div(tiles, 256))), dim3(256), 0, stream,
                         int32_t(tiles), ws->score, ws->counts + kFixedIdx);
      a.w.fixed = ws->score;
    } else {
      const uint32_t* trig = io->trigger ? io->trigger : io->seeds;
      hipLaunchKernelGGL(ks.init_ids, dim3(unsigned(ceil_div(tiles, 64))), dim3(64), 0, stream, trig,
                         int32_t(tiles), d->activate_all, ws->stamp, ws->lists, ws->counts,
                         a.w.shard_cap);
    }
    return AGD_OK;
  };
  auto launch = [&](int32_t koff, hipStream_t s) {
    a.w.koff = koff;
    hipLaunchKernelGGL(ks.update_batch, dim3(blocks), dim3(ni), 0, s, a);
  };
  WorklistResult r;
  int st = run_worklist(int32_t(tiles), blocks, d->nitermax_o, d->check_interval, stream, init,
                        launch, r, d->time_launches != 0, d->use_graph != 0);
  if (st) return st;
  if (stats) {
    stats->kernel_ms = r.kernel_ms;
    stats->niter_o = r.niter;
    stats->tile_updates = r.visits;
    stats->cell_updates = r.visits * ni;
    stats->launches = r.passes;
    stats->gpu_ms = r.ms;
    stats->converged = r.converged ? 1 : 0;
    stats->grid_blocks = blocks;
  }
  if (!r.converged) {
    set_error("batched eikonal solve did not converge within nitermax_o passes");
    return AGD_NOCONV;
  }
  return AGD_OK;
}

template <class S>
static int scheme_impl(const agd_eikonal_desc* d, const agd_eikonal_io* io, hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(ks.scheme, "no precomputed-scheme variant for this model");
  AGD_REQUIRE(io && io->geom && io->scheme_weights && io->scheme_codes,
              "geom, scheme_weights and scheme_codes are required");
  const GridDesc g = make_grid(d);
  SchemeArgs<S> a{};
  a.geom = static_cast<const S*>(io->geom);
  a.wall = io->wall_dist;
  a.g = g;
  for (int k = 0; k < 3; ++k) a.iso_w[k] = S(d->iso_weights[k]);
  a.sw = static_cast<S*>(io->scheme_weights);
  a.scode = io->scheme_codes;
  hipLaunchKernelGGL((io->wall_dist || g.periodic) ? ks.scheme_general : ks.scheme, dim3(g.ntiles),
                     dim3(ks.ni), 0, stream, a);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

}  // namespace agd

using namespace agd;

extern "C" {

int agd_eikonal_solve_batched(const agd_eikonal_desc* desc, int32_t ninst,
                              const agd_batch_instance* instances, const agd_eikonal_io* io,
                              agd_eikonal_stats* stats, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32)
    return solve_batched_impl<float>(desc, ninst, instances, io, stats, as_stream(stream));
  return solve_batched_impl<double>(desc, ninst, instances, io, stats, as_stream(stream));
}

int agd_scheme_supported(int32_t model, int32_t ndim, const int32_t* shape_i) {
  KernelSet<float> ks;
  return (shape_i && lookup<float>(model, ndim, shape_i, ks) && ks.scheme) ? 1 : 0;
}

int agd_eikonal_scheme(const agd_eikonal_desc* desc, const agd_eikonal_io* io, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32) return scheme_impl<float>(desc, io, as_stream(stream));
  return scheme_impl<double>(desc, io, as_stream(stream));
}

int agd_tile_supported(int32_t model, int32_t ndim, const int32_t* shape_i) {
  KernelSet<float> ks;
  return (shape_i && lookup<float>(model, ndim, shape_i, ks)) ? 1 : 0;
}

int agd_eikonal_solve(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                      agd_eikonal_stats* stats, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32) return solve_impl<float>(desc, io, stats, as_stream(stream));
  return solve_impl<double>(desc, io, stats, as_stream(stream));
}

int agd_eikonal_flow(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                     const agd_flow_io* flow, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32) return flow_impl<float>(desc, io, flow, as_stream(stream));
  return flow_impl<double>(desc, io, flow, as_stream(stream));
}

int agd_stencil_export(int32_t model, int32_t ndim, int32_t scalar, int64_t n, const void* geom,
                       const double* iso_weights, void* weights, int32_t* offsets, void* stream) {
  AGD_REQUIRE(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
  AGD_REQUIRE(n >= 0 && weights && offsets, "bad buffers");
  AGD_REQUIRE(model == AGD_ISOTROPIC || geom, "geometry required");
  if (n == 0) return AGD_OK;
  const double iw[3] = {iso_weights ? iso_weights[0] : 0., iso_weights ? iso_weights[1] : 0.,
                        iso_weights ? iso_weights[2] : 0.};
  hipStream_t s = as_stream(stream);
  const dim3 grid(unsigned(ceil_div(n, 256))), block(256);
#define AGD_EXPORT(S, M, D)                                                                  \
  hipLaunchKernelGGL((stencil_export_kernel<S, M, D>), grid, block, 0, s, n,                 \
                     static_cast<const S*>(geom), S(iw[0]), S(iw[1]), S(iw[2]),              \
                     static_cast<S*>(weights), offsets)
#define AGD_EXPORT_M(S)                                                         \
  switch (model * 10 + ndim) {                                                  \
    case AGD_ISOTROPIC * 10 + 2: AGD_EXPORT(S, AGD_ISOTROPIC, 2); break;        \
    case AGD_ISOTROPIC * 10 + 3: AGD_EXPORT(S, AGD_ISOTROPIC, 3); break;        \
    case AGD_DIAGONAL * 10 + 2: AGD_EXPORT(S, AGD_DIAGONAL, 2); break;          \
    case AGD_DIAGONAL * 10 + 3: AGD_EXPORT(S, AGD_DIAGONAL, 3); break;          \
    case AGD_RIEMANN * 10 + 2: AGD_EXPORT(S, AGD_RIEMANN, 2); break;            \
    case AGD_RIEMANN * 10 + 3: AGD_EXPORT(S, AGD_RIEMANN, 3); break;            \
    case AGD_RANDER * 10 + 2: AGD_EXPORT(S, AGD_RANDER, 2); break;              \
    case AGD_RANDER * 10 + 3: AGD_EXPORT(S, AGD_RANDER, 3); break;              \
    default: set_error("invalid argument: unknown model"); return AGD_EINVAL;   \
  }
  if (scalar == AGD_F32) { AGD_EXPORT_M(float) }
  else { AGD_EXPORT_M(double) }
#undef AGD_EXPORT_M
#undef AGD_EXPORT
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_scatter_seeds(int32_t scalar, int64_t nseeds, const int64_t* tiled_index,
                      const double* seed_values, void* rhs, uint32_t* seeds, uint32_t* trigger,
                      void* stream) {
  AGD_REQUIRE(nseeds >= 0, "negative seed count");
  if (nseeds == 0) return AGD_OK;
  AGD_REQUIRE(tiled_index && seed_values && rhs && seeds, "null buffer");
  hipStream_t s = as_stream(stream);
  const dim3 grid(unsigned(ceil_div(nseeds, 256))), block(256);
  if (scalar == AGD_F32)
    hipLaunchKernelGGL(scatter_seeds_kernel<float>, grid, block, 0, s, nseeds, tiled_index,
                       seed_values, static_cast<float*>(rhs), seeds, trigger);
  else
    hipLaunchKernelGGL(scatter_seeds_kernel<double>, grid, block, 0, s, nseeds, tiled_index,
                       seed_values, static_cast<double*>(rhs), seeds, trigger);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_prepare_walls(const agd_eikonal_desc* desc, const uint8_t* walls, int32_t bound,
                      void* rhs, uint32_t* seeds, uint8_t* wall_dist, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  AGD_REQUIRE(walls && rhs && seeds && wall_dist, "null buffer");
  AGD_REQUIRE(bound >= 0 && bound < 255, "wallDistBound must lie in [0,255)");
  KernelSet<float> ks;
  AGD_REQUIRE(lookup<float>(desc->model, desc->ndim, desc->shape_i, ks), "unsupported tile shape");
  const GridDesc g = make_grid(desc);
  hipStream_t s = as_stream(stream);
  const int64_t N = g.size_tot;
  const dim3 grid(unsigned(ceil_div(N, 256))), block(256);
  if (desc->scalar == AGD_F32)
    hipLaunchKernelGGL(walls_mark_kernel<float>, grid, block, 0, s, N, walls,
                       static_cast<float*>(rhs), seeds, wall_dist);
  else
    hipLaunchKernelGGL(walls_mark_kernel<double>, grid, block, 0, s, N, walls,
                       static_cast<double*>(rhs), seeds, wall_dist);
  AGD_HIP_TRY(hipGetLastError());
  if (bound > 0) {
    uint8_t* tmp = nullptr;
    AGD_HIP_TRY(hipMalloc((void**)&tmp, N));
    uint8_t* src = wall_dist;
    uint8_t* dst = tmp;
    for (int i = 0; i < bound; ++i) {
      hipLaunchKernelGGL(ks.wallpass, dim3(g.ntiles), dim3(ks.ni), 0, s, g, src, dst);
      std::swap(src, dst);
    }
    if (src != wall_dist)
      AGD_HIP_TRY(hipMemcpyAsync(wall_dist, src, N, hipMemcpyDeviceToDevice, s));
    AGD_HIP_TRY(hipStreamSynchronize(s));
    AGD_HIP_TRY(hipFree(tmp));
  }
  hipLaunchKernelGGL(wall_dist_cap, grid, block, 0, s, N, wall_dist, bound);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

}  // extern "C"
