// device.hpp -- device building blocks of the eikonal tile solver (gfx950).
//
//  * tiled (bi-level) indexing: n_t = n_o * size_i + n_i, tile shape known at
//    compile time so that in-tile divisions become shifts;
//  * Selling decomposition of 2x2 / 3x3 symmetric positive definite matrices,
//    kept entirely in registers (no dynamically indexed private arrays: the
//    reduction loop is unrolled over its cyclic phases so every superbase
//    access has a compile-time index);
//  * the upwind Hopf-Lax update: sorting network on (value, weight) pairs and
//    the active-set quadratic solve;
//  * the wall visibility walk.
//
// Semantics follow the reference device code (paths relative to the AGD
// source tree): Geometry_.h:144-156 (scal_vmv), Geometry2.h:21-48,
// Geometry3.h:19-58 (Selling), Isotropic_.h:20-41 and Riemann_.h:16-29
// (schemes), HFM.h:28-207 (update), NetworkSort.h:17-60 (sort networks),
// Walls.h:13-50 (visibility), Grid.h:15-90 (indexing).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/agd_hip.h"

namespace agd {

template <class S> __device__ __host__ inline S inf_v();
template <> __device__ __host__ inline float inf_v<float>() { return __builtin_huge_valf(); }
template <> __device__ __host__ inline double inf_v<double>() { return __builtin_huge_val(); }

__device__ inline float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float abs_(float a) { return __builtin_fabsf(a); }
__device__ inline double abs_(double a) { return __builtin_fabs(a); }
__device__ inline float sqrt_(float a) { return __builtin_sqrtf(a); }
__device__ inline double sqrt_(double a) { return __builtin_sqrt(a); }
// Square root and quotient of the update: the hardware instructions in FP32
// (v_sqrt_f32, v_rcp_f32: about 1 ulp, against 17 and 11 instructions for
// the correctly rounded sequences; the update's error budget is 1e-4 of the
// values), correctly rounded in FP64.
__device__ inline float sqrt_fast(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ inline double sqrt_fast(double a) { return __builtin_sqrt(a); }
__device__ inline float div_fast(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline double div_fast(double a, double b) { return a / b; }
__device__ inline bool isnan_(float a) { return a != a; }
__device__ inline bool isnan_(double a) { return a != a; }

// Product of two integers of magnitude below 2^23 whose product fits in 32
// bits: a full-rate 24-bit multiply (v_mul_i32_i24) where a 32-bit one
// (v_mul_lo_u32) issues at quarter rate.  Used for tile coordinates and
// extents (below 2^15), and for superbase entries and stencil offsets,
// which stay below sqrt(6 cond(D)) in magnitude (the reduction decreases
// sum_i |b_i|_D^2), far below 2^23 for any FP32 metric.
__device__ inline int imul24(int a, int b) { return __mul24(a, b); }

// base[idx] for an array spanning less than 4 GiB: the byte offset is formed
// in 32 bits, so that the access takes the scalar-base + 32-bit-offset form
// (one address VGPR per gather instead of a 64-bit pair).
template <class T> __device__ inline T ld32(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * uint32_t(sizeof(T)));
}
template <class T> __device__ inline void st32(T* base, uint32_t idx, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * uint32_t(sizeof(T))) = v;
}

// ------------------------------------------------------------------ models
template <int MODEL, int D> struct ModelTraits {
  static constexpr int SYMDIM = D * (D + 1) / 2;
  // Curvature-penalised models (3D grid (x, y, theta)).
  static constexpr bool CURV = MODEL >= AGD_REEDSSHEPP2 && MODEL <= AGD_DUBINS2;
  // Offsets of one sub-scheme, each a pair of neighbours x -+ e (the last
  // NFWD of them forward: x + e only, the x - e side left out).
  static constexpr int NSYM = (MODEL == AGD_ISOTROPIC || MODEL == AGD_DIAGONAL) ? D
                              : MODEL == AGD_REEDSSHEPP2                     ? 6
                              : MODEL == AGD_REEDSSHEPPFWD2                  ? 7
                              : MODEL == AGD_ELASTICA2                       ? 30
                              : MODEL == AGD_DUBINS2                         ? 6
                                                                             : SYMDIM;
  static constexpr int NFWD = MODEL == AGD_REEDSSHEPPFWD2 ? 6
                              : MODEL == AGD_ELASTICA2    ? 30
                              : MODEL == AGD_DUBINS2      ? 6
                                                          : 0;
  static constexpr int NGEOM = MODEL == AGD_ISOTROPIC ? 0
                               : MODEL == AGD_DIAGONAL ? D
                               : MODEL == AGD_RIEMANN  ? SYMDIM
                               : MODEL == AGD_RANDER   ? SYMDIM + D
                               : CURV                  ? 4
                                                       : 5 + D * D;
  static constexpr bool DRIFT = (MODEL == AGD_RANDER);
  // TTI: the update is the min or max over NMIX Selling schemes; Dubins2 the
  // min over its left and right turning schemes.
  static constexpr int NMIX = MODEL == AGD_TTI ? 7 : (MODEL == AGD_DUBINS2 ? 2 : 1);
  // The scheme is always precomputed (agd_eikonal_scheme) and read per sweep.
  static constexpr bool STORED = MODEL == AGD_TTI || CURV;
};

// ------------------------------------------------------------------ tiles
template <int D, int T0, int T1, int T2> struct Tile {
  static constexpr int NI = T0 * T1 * T2;
  static constexpr int T[3] = {T0, T1, T2};
  // Row-major position of cell n inside the tile (last axis fastest).
  __device__ static inline void pos(int n, int* x) {
    if constexpr (D == 2) {
      x[1] = n % T1;
      x[0] = n / T1;
      x[2] = 0;
    } else {
      x[2] = n % T2;
      x[1] = (n / T2) % T1;
      x[0] = n / (T1 * T2);
    }
  }
  __device__ static inline int index(const int* x) {
    if constexpr (D == 2) return x[0] * T1 + x[1];
    else return (x[0] * T1 + x[1]) * T2 + x[2];
  }
  __device__ static inline bool inside(const int* x) {
    bool r = (x[0] >= 0) && (x[0] < T0) && (x[1] >= 0) && (x[1] < T1);
    if constexpr (D == 3) r = r && (x[2] >= 0) && (x[2] < T2);
    return r;
  }
};

// Grid-level description passed by value to every kernel.
struct GridDesc {
  int32_t shape_tot[3];  // grid shape
  int32_t shape_o[3];    // tile grid shape
  int32_t periodic;      // bit k: axis k periodic
  int32_t ntiles;
  int64_t size_tot;      // ntiles * size_i
};

__device__ inline int mod_pos(int x, int n) {
  x %= n;
  return x >= 0 ? x : x + n;
}

// Worklist entries are packed tile coordinates (10 bits per axis in 3D, 15
// in 2D), so a pass decodes its tile with shifts instead of divisions.
template <int D> __device__ __host__ inline int pack_tile(const int* xo) {
  if constexpr (D == 2) return xo[0] | (xo[1] << 15);
  else return xo[0] | (xo[1] << 10) | (xo[2] << 20);
}
template <int D> __device__ inline void unpack_tile(int p, int* xo) {
  if constexpr (D == 2) {
    xo[0] = p & 0x7fff;
    xo[1] = (p >> 15) & 0x7fff;
    xo[2] = 0;
  } else {
    xo[0] = p & 0x3ff;
    xo[1] = (p >> 10) & 0x3ff;
    xo[2] = (p >> 20) & 0x3ff;
  }
}
template <int D> __device__ inline int tile_linear(const int* xo, const GridDesc& g) {
  if constexpr (D == 2) return imul24(xo[0], g.shape_o[1]) + xo[1];
  else return imul24(imul24(xo[0], g.shape_o[1]) + xo[1], g.shape_o[2]) + xo[2];
}

// Position of tile n_o in the tile grid.
template <int D> __device__ inline void tile_pos(int n_o, const GridDesc& g, int* xo) {
  if constexpr (D == 2) {
    xo[1] = n_o % g.shape_o[1];
    xo[0] = n_o / g.shape_o[1];
    xo[2] = 0;
  } else {
    xo[2] = n_o % g.shape_o[2];
    const int r = n_o / g.shape_o[2];
    xo[1] = r % g.shape_o[1];
    xo[0] = r / g.shape_o[1];
  }
}

// Grid.h:68-90 (InRange_per + Index_per on the tile grid): index of the
// neighbouring tile of xo along axis `ax` in direction `dir`, -1 if outside.
// (Axis selection by comparison keeps every array access compile-time.)
template <int D>
__device__ inline int tile_neighbor(const int* xo, int ax, int dir, const GridDesc& g,
                                    int* packed = nullptr) {
  int y[3] = {xo[0], xo[1], xo[2]};
  bool ok = true;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (k == ax) {
      y[k] += dir;
      if ((g.periodic >> k) & 1) y[k] = mod_pos(y[k], g.shape_o[k]);
      else ok = y[k] >= 0 && y[k] < g.shape_o[k];
    }
  }
  if (!ok) return -1;
  if (packed) *packed = pack_tile<D>(y);
  return tile_linear<D>(y, g);
}

// Grid.h:15-35 Index_tot after the InRange_per test: tiled linear index of
// the global position y (wrapped along periodic axes when PER), or -1
// outside.  Tiled indices fit in 32 bits (the host rejects grids of 2^31
// cells or more).
template <int D, class TL, bool PER = true>
__device__ inline int index_tot(const int* y_in, const GridDesc& g) {
  int y[3] = {y_in[0], y_in[1], y_in[2]};
  bool ok = true;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (PER && ((g.periodic >> k) & 1)) y[k] = mod_pos(y[k], g.shape_tot[k]);
    else ok = ok && y[k] >= 0 && y[k] < g.shape_tot[k];
  }
  if (!ok) return -1;
  int yo[3], yi[3];
#pragma unroll
  for (int k = 0; k < D; ++k) {  // y >= 0 here: unsigned division by a power of two
    yo[k] = int(unsigned(y[k]) / unsigned(TL::T[k]));
    yi[k] = int(unsigned(y[k]) % unsigned(TL::T[k]));
  }
  if constexpr (D == 2) {
    return (imul24(yo[0], g.shape_o[1]) + yo[1]) * TL::NI + (yi[0] * TL::T[1] + yi[1]);
  } else {
    return (imul24(imul24(yo[0], g.shape_o[1]) + yo[1], g.shape_o[2]) + yo[2]) * TL::NI +
           ((yi[0] * TL::T[1] + yi[1]) * TL::T[2] + yi[2]);
  }
}

// ------------------------------------------------------------------ Selling
// Geometry_.h:144-156: integer products first, then one (fused) multiply-add
// per coefficient of the lower triangle (d00, d10, d11, d20, d21, d22).
template <class S> __device__ inline S scal_vmv2(const int* x, const S* m, const int* y) {
  S r = S(0);
  r = fma_(S(imul24(x[0], y[0])), m[0], r);
  r = fma_(S(imul24(x[1], y[0]) + imul24(x[0], y[1])), m[1], r);
  r = fma_(S(imul24(x[1], y[1])), m[2], r);
  return r;
}
template <class S> __device__ inline S scal_vmv3(const int* x, const S* m, const int* y) {
  S r = S(0);
  r = fma_(S(imul24(x[0], y[0])), m[0], r);
  r = fma_(S(imul24(x[1], y[0]) + imul24(x[0], y[1])), m[1], r);
  r = fma_(S(imul24(x[1], y[1])), m[2], r);
  r = fma_(S(imul24(x[2], y[0]) + imul24(x[0], y[2])), m[3], r);
  r = fma_(S(imul24(x[2], y[1]) + imul24(x[1], y[2])), m[4], r);
  r = fma_(S(imul24(x[2], y[2])), m[5], r);
  return r;
}

constexpr int kSellingMaxIter = 100;

// One cyclic step (i,j,k) of the 2D reduction (Geometry2.h:25-33).
template <class S, int I, int J, int K>
__device__ inline void selling2_step(int (&sb)[3][2], const S* m, int& red) {
  if (scal_vmv2<S>(sb[I], m, sb[J]) > S(0)) {
    sb[K][0] = sb[I][0] - sb[J][0];
    sb[K][1] = sb[I][1] - sb[J][1];
    sb[I][0] = -sb[I][0];
    sb[I][1] = -sb[I][1];
    red = 0;
  }
}

// Selling decomposition, 2D, reference order of operations: weights w[3],
// offsets e[3][2].
template <class S> __device__ inline void selling2_int(const S* m, S* w, int (&e)[3][3]) {
  int sb[3][2] = {{1, 0}, {0, 1}, {-1, -1}};
  int it = 0, red = 0;
  bool done = false;
  for (int outer = 0; outer < (kSellingMaxIter + 2) / 3 + 1 && !done; ++outer) {
    // Phases 0,1,2 of iter%3 with compile-time indices.
    if (!done) { if (it < kSellingMaxIter && red < 3) { selling2_step<S, 0, 1, 2>(sb, m, red); ++it; ++red; } else done = true; }
    if (!done) { if (it < kSellingMaxIter && red < 3) { selling2_step<S, 1, 2, 0>(sb, m, red); ++it; ++red; } else done = true; }
    if (!done) { if (it < kSellingMaxIter && red < 3) { selling2_step<S, 2, 0, 1>(sb, m, red); ++it; ++red; } else done = true; }
  }
  // Geometry2.h:40-48: w_r = max(0, -<b_r, D b_{r+1}>), e_r = perp(b_{r+2}).
  const S s0 = -scal_vmv2<S>(sb[0], m, sb[1]);
  const S s1 = -scal_vmv2<S>(sb[1], m, sb[2]);
  const S s2 = -scal_vmv2<S>(sb[2], m, sb[0]);
  w[0] = s0 > S(0) ? s0 : S(0);
  w[1] = s1 > S(0) ? s1 : S(0);
  w[2] = s2 > S(0) ? s2 : S(0);
  e[0][0] = -sb[2][1]; e[0][1] = sb[2][0]; e[0][2] = 0;
  e[1][0] = -sb[0][1]; e[1][1] = sb[0][0]; e[1][2] = 0;
  e[2][0] = -sb[1][1]; e[2][1] = sb[1][0]; e[2][2] = 0;
}

// One step of the 3D reduction with permutation (i,j,k,l) (Geometry3.h:27-37),
// predicated: the step takes effect only in lanes still iterating (`live`),
// which keeps the loop free of divergent branches (and of the registers
// their merges cost).
template <class S, int I, int J, int K, int L>
__device__ inline void selling3_step(int (&sb)[4][3], const S* m, int& it, int& red) {
  const bool live = it < kSellingMaxIter && red < 6;
  const bool acute = scal_vmv3<S>(sb[I], m, sb[J]) > S(0);
  const bool flip = live && acute;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int bi = sb[I][c];
    sb[K][c] = flip ? sb[K][c] + bi : sb[K][c];
    sb[L][c] = flip ? sb[L][c] + bi : sb[L][c];
    sb[I][c] = flip ? -bi : bi;
  }
  red = flip ? 1 : (live ? red + 1 : red);
  it = live ? it + 1 : it;
}

template <int I, int J, int K, int L, class S>
__device__ inline void selling3_out(const int (&sb)[4][3], const S* m, S& w, int* e) {
  const S s = -scal_vmv3<S>(sb[I], m, sb[J]);
  w = s > S(0) ? s : S(0);
  e[0] = imul24(sb[K][1], sb[L][2]) - imul24(sb[K][2], sb[L][1]);
  e[1] = imul24(sb[K][2], sb[L][0]) - imul24(sb[K][0], sb[L][2]);
  e[2] = imul24(sb[K][0], sb[L][1]) - imul24(sb[K][1], sb[L][0]);
}

// Selling decomposition, 3D, reference order of operations: weights w[6],
// offsets e[6][3].  Permutations {0,1,2,3},{0,2,1,3},{0,3,1,2},{1,2,0,3},
// {1,3,0,2},{2,3,0,1}.  The loop runs in rounds of the 6 phases until no lane
// of the wave is still iterating.
template <class S> __device__ inline void selling3_int(const S* m, S* w, int (&e)[6][3]) {
  int sb[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, -1, -1}};
  int it = 0, red = 0;
#pragma unroll 1
  for (int round = 0; round < (kSellingMaxIter + 5) / 6; ++round) {
    selling3_step<S, 0, 1, 2, 3>(sb, m, it, red);
    selling3_step<S, 0, 2, 1, 3>(sb, m, it, red);
    selling3_step<S, 0, 3, 1, 2>(sb, m, it, red);
    selling3_step<S, 1, 2, 0, 3>(sb, m, it, red);
    selling3_step<S, 1, 3, 0, 2>(sb, m, it, red);
    selling3_step<S, 2, 3, 0, 1>(sb, m, it, red);
    if (!__any(it < kSellingMaxIter && red < 6)) break;
  }
  selling3_out<0, 1, 2, 3>(sb, m, w[0], e[0]);
  selling3_out<0, 2, 1, 3>(sb, m, w[1], e[1]);
  selling3_out<0, 3, 1, 2>(sb, m, w[2], e[2]);
  selling3_out<1, 2, 0, 3>(sb, m, w[3], e[3]);
  selling3_out<1, 3, 0, 2>(sb, m, w[4], e[4]);
  selling3_out<2, 3, 0, 1>(sb, m, w[5], e[5]);
}

// <x, m y> with the superbase entries held as floats.  While every entry is
// an integer of magnitude <= kSellingExactEntry = 2^11, every product and
// sum of two products below is an integer below 2^24, exact in FP32: the
// coefficient matches the integer one of scal_vmv3 and the fused
// multiply-adds that follow are the same operations in the same order.  The
// result is bit-identical, without the six integer-to-float conversions.
constexpr float kSellingExactEntry = 2048.f;
template <class S> __device__ inline S scal_vmv3f(const float* x, const S* m, const float* y) {
  S r = S(0);
  r = fma_(S(x[0] * y[0]), m[0], r);
  r = fma_(S(fma_(x[1], y[0], x[0] * y[1])), m[1], r);
  r = fma_(S(x[1] * y[1]), m[2], r);
  r = fma_(S(fma_(x[2], y[0], x[0] * y[2])), m[3], r);
  r = fma_(S(fma_(x[2], y[1], x[1] * y[2])), m[4], r);
  r = fma_(S(x[2] * y[2]), m[5], r);
  return r;
}

// One phase (i,j,k,l) of the 3D reduction (Geometry3.h:27-37).  A lane is
// still iterating while red < 6 (phases since its last flip) and the phase
// index is below kSellingMaxIter: lanes that stopped have the same count of
// phases as every lane still iterating, so the iteration cap is one
// wave-uniform flag.  The flip itself runs only when some lane of the wave
// flips (most phases of a tile flip in no lane: neighbouring cells have
// close metrics).  pr[P] keeps <b_i, m b_j> of the current superbase: a flip
// negates b_i and leaves b_j, and rounding to nearest is symmetric, so the
// product after the flip is exactly the negated one.
template <class S, int I, int J, int K, int L, int P>
__device__ inline void selling3f_step(float (&sb)[4][3], const S* m, S (&pr)[6], int& red,
                                      float& emax, bool below_cap) {
  const S r = scal_vmv3f<S>(sb[I], m, sb[J]);
  const bool flip = below_cap && red < 6 && r > S(0);
  red = flip ? 1 : red + 1;
  pr[P] = flip ? -r : r;
  if (__any(flip)) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float bi = sb[I][c];
      sb[K][c] = flip ? sb[K][c] + bi : sb[K][c];
      sb[L][c] = flip ? sb[L][c] + bi : sb[L][c];
      sb[I][c] = flip ? -bi : bi;
    }
    // a flip only enlarges b_k and b_l
    const float ek = __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(sb[K][0]), __builtin_fabsf(sb[K][1])),
                                     __builtin_fabsf(sb[K][2]));
    const float el = __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(sb[L][0]), __builtin_fabsf(sb[L][1])),
                                     __builtin_fabsf(sb[L][2]));
    emax = __builtin_fmaxf(emax, __builtin_fmaxf(ek, el));
  }
}

template <int I, int J, int K, int L, class S>
__device__ inline void selling3f_out(const int (&sb)[4][3], S p, S& w, int* e) {
  const S s = -p;
  w = s > S(0) ? s : S(0);
  e[0] = imul24(sb[K][1], sb[L][2]) - imul24(sb[K][2], sb[L][1]);
  e[1] = imul24(sb[K][2], sb[L][0]) - imul24(sb[K][0], sb[L][2]);
  e[2] = imul24(sb[K][0], sb[L][1]) - imul24(sb[K][1], sb[L][0]);
}

// Selling decomposition, 3D: weights w[6], offsets e[6][3], bit-identical to
// selling3_int.  The weights are the products of the last phase of each pair
// (after the last flip, the six phases that end the reduction see the final
// superbase).  A wave in which some lane stopped at the iteration cap, or
// had a superbase entry beyond kSellingExactEntry (metrics of condition
// number above about 7e5), redoes the reduction in integers.
template <class S> __device__ inline void selling3(const S* m, S* w, int (&e)[6][3]) {
  float sb[4][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}, {-1.f, -1.f, -1.f}};
  S pr[6];
  int red = 0;
  float emax = 1.f;
#pragma unroll 1
  for (int step = 0; step < kSellingMaxIter; step += 6) {
    selling3f_step<S, 0, 1, 2, 3, 0>(sb, m, pr, red, emax, step + 0 < kSellingMaxIter);
    selling3f_step<S, 0, 2, 1, 3, 1>(sb, m, pr, red, emax, step + 1 < kSellingMaxIter);
    selling3f_step<S, 0, 3, 1, 2, 2>(sb, m, pr, red, emax, step + 2 < kSellingMaxIter);
    selling3f_step<S, 1, 2, 0, 3, 3>(sb, m, pr, red, emax, step + 3 < kSellingMaxIter);
    selling3f_step<S, 1, 3, 0, 2, 4>(sb, m, pr, red, emax, step + 4 < kSellingMaxIter);
    selling3f_step<S, 2, 3, 0, 1, 5>(sb, m, pr, red, emax, step + 5 < kSellingMaxIter);
    if (!__any(red < 6)) break;
  }
  if (__any(red < 6 || emax > kSellingExactEntry)) {
    selling3_int<S>(m, w, e);
    return;
  }
  int ib[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) ib[i][c] = int(sb[i][c]);
  selling3f_out<0, 1, 2, 3>(ib, pr[0], w[0], e[0]);
  selling3f_out<0, 2, 1, 3>(ib, pr[1], w[1], e[1]);
  selling3f_out<0, 3, 1, 2>(ib, pr[2], w[2], e[2]);
  selling3f_out<1, 2, 0, 3>(ib, pr[3], w[3], e[3]);
  selling3f_out<1, 3, 0, 2>(ib, pr[4], w[4], e[4]);
  selling3f_out<2, 3, 0, 1>(ib, pr[5], w[5], e[5]);
}

// 2D counterparts of the above: <x, m y> on float-held integer entries
// (bit-identical to scal_vmv2 while entries stay within kSellingExactEntry)
// and one cyclic phase (i,j,k) of Geometry2.h:25-33.  Pair (i, i+1 mod 3)
// is the pair of weight i; a flip negates b_i and leaves b_j, so pr[I]
// becomes exactly the negated product.
template <class S> __device__ inline S scal_vmv2f(const float* x, const S* m, const float* y) {
  S r = S(0);
  r = fma_(S(x[0] * y[0]), m[0], r);
  r = fma_(S(fma_(x[1], y[0], x[0] * y[1])), m[1], r);
  r = fma_(S(x[1] * y[1]), m[2], r);
  return r;
}

template <class S, int I, int J, int K>
__device__ inline void selling2f_step(float (&sb)[3][2], const S* m, S (&pr)[3], int& red,
                                      float& emax, bool below_cap) {
  const S r = scal_vmv2f<S>(sb[I], m, sb[J]);
  const bool flip = below_cap && red < 3 && r > S(0);
  red = flip ? 1 : red + 1;
  pr[I] = flip ? -r : r;
  if (__any(flip)) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const float bi = sb[I][c];
      sb[K][c] = flip ? bi - sb[J][c] : sb[K][c];
      sb[I][c] = flip ? -bi : bi;
    }
    emax = __builtin_fmaxf(emax, __builtin_fmaxf(__builtin_fabsf(sb[K][0]), __builtin_fabsf(sb[K][1])));
  }
}

// Selling decomposition, 2D: weights w[3], offsets e[3][2], bit-identical to
// selling2_int (same scheme as selling3: weights from the last phase of each
// pair, integer path for a wave with a lane at the iteration cap or beyond
// kSellingExactEntry).
template <class S> __device__ inline void selling2(const S* m, S* w, int (&e)[3][3]) {
  float sb[3][2] = {{1.f, 0.f}, {0.f, 1.f}, {-1.f, -1.f}};
  S pr[3];
  int red = 0;
  float emax = 1.f;
#pragma unroll 1
  for (int step = 0; step < kSellingMaxIter; step += 3) {
    selling2f_step<S, 0, 1, 2>(sb, m, pr, red, emax, step + 0 < kSellingMaxIter);
    selling2f_step<S, 1, 2, 0>(sb, m, pr, red, emax, step + 1 < kSellingMaxIter);
    selling2f_step<S, 2, 0, 1>(sb, m, pr, red, emax, step + 2 < kSellingMaxIter);
    if (!__any(red < 3)) break;
  }
  if (__any(red < 3 || emax > kSellingExactEntry)) {
    selling2_int<S>(m, w, e);
    return;
  }
  // Geometry2.h:40-48: w_r = max(0, -<b_r, D b_{r+1}>), e_r = perp(b_{r+2}).
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const S sr = -pr[r];
    w[r] = sr > S(0) ? sr : S(0);
  }
  e[0][0] = -int(sb[2][1]); e[0][1] = int(sb[2][0]); e[0][2] = 0;
  e[1][0] = -int(sb[0][1]); e[1][1] = int(sb[0][0]); e[1][2] = 0;
  e[2][0] = -int(sb[1][1]); e[2][1] = int(sb[1][0]); e[2][2] = 0;
}

// Per-cell stencil: NSYM symmetric offsets +-e_k with weights w_k, plus the
// drift shift <omega, e_k> for Rander (FiniteDifferences.h:41-45).
template <class S, int MODEL, int D> struct Stencil {
  using MT = ModelTraits<MODEL, D>;
  S w[MT::NSYM];
  int e[MT::NSYM][3];
  S shift[MT::NSYM];
};

template <class S, int MODEL, int D>
__device__ inline void make_stencil(const S* g, const S* iso_w, Stencil<S, MODEL, D>& st) {
  using MT = ModelTraits<MODEL, D>;
  if constexpr (MODEL == AGD_ISOTROPIC || MODEL == AGD_DIAGONAL) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      st.w[k] = (MODEL == AGD_ISOTROPIC) ? iso_w[k] : g[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) st.e[k][c] = (c == k);
      st.shift[k] = S(0);
    }
  } else {
    if constexpr (D == 2) selling2<S>(g, st.w, st.e);
    else selling3<S>(g, st.w, st.e);
#pragma unroll
    for (int k = 0; k < MT::NSYM; ++k) {
      S s = S(0);
      if constexpr (MT::DRIFT) {
#pragma unroll
        for (int c = 0; c < D; ++c) s += g[MT::SYMDIM + c] * S(st.e[k][c]);
      }
      st.shift[k] = s;
    }
  }
}

// ------------------------------------------------------------------ TTI
// TTI_.h:41-134: the dual unit ball <l, x> + 1/2 <x, q x> = 1 in the squared
// coordinates x = (X^2 [+ Y^2], Z^2) of T p is replaced by NMIX tangent
// ellipses; ellipse i has the dual metric diag_i[0] D0 + diag_i[1] D1 with D0,
// D1 the outer products of T's rows.  Written without contraction (each
// product rounded), as the oracle, so FP32 results agree bit for bit.
constexpr int kTTIMix = 7;

template <class S> __device__ inline void tti_solve2(S a, S b, S c, S (&r)[2]) {
#pragma clang fp contract(off)
  const S sdelta = sqrt_(b * b - a * c);
  const S u = -b + sdelta, v = -sdelta - b;
  if (abs_(c) > abs_(a)) {
    r[0] = u / c;
    r[1] = v / c;
  } else if (a != S(0)) {
    r[0] = a / u;
    r[1] = a / v;
  } else {
    r[0] = S(0);
    r[1] = inf_v<S>();
  }
}

template <class S> __device__ inline S tti_solve2_above(S a, S b, S c, S above) {
  S r[2];
  tti_solve2<S>(a, b, c, r);
  const bool ordered = r[0] < r[1];
  const S rmin = ordered ? r[0] : r[1], rmax = ordered ? r[1] : r[0];
  return rmin >= above ? rmin : rmax;
}

template <class S>
__device__ inline void tti_grad_ratio(const S* l, const S* q, S x0, S x1, S& g0, S& g1) {
#pragma clang fp contract(off)
  S h0 = q[0] * x0 + q[1] * x1;
  S h1 = q[1] * x0 + q[2] * x1;
  h0 = l[0] + h0;
  h1 = l[1] + h1;
  const S xhg = x0 * h0 + x1 * h1;
  g0 = h0 / xhg;
  g1 = h1 / xhg;
}

template <class S> __device__ inline S tti_vqv(S x0, S x1, const S* q, S y0, S y1) {
#pragma clang fp contract(off)
  return x0 * q[0] * y0 + (x0 * y1 + x1 * y0) * q[1] + x1 * q[2] * y1;
}

// diags(): diag[i] = grad f / <x, grad f> at NMIX points of the curve; returns
// mix_is_min = det(diag[0], diag[NMIX-1]) > 0.
template <class S> __device__ inline bool tti_diags(const S* l, const S* q, S (&diag)[kTTIMix][2]) {
#pragma clang fp contract(off)
  const S a = tti_solve2_above<S>(S(-2), l[0], q[0], S(0));
  const S b = tti_solve2_above<S>(S(-2), l[1], q[2], S(0));
  const S L0 = l[0] * S(0.5) + l[1] * S(0.5), L1 = l[0] * S(0.5) + l[1] * S(-0.5);
  const S Q0 = tti_vqv<S>(S(0.5), S(0.5), q, S(0.5), S(0.5));
  const S Q1 = tti_vqv<S>(S(0.5), S(0.5), q, S(0.5), S(-0.5));
  const S Q2 = tti_vqv<S>(S(0.5), S(-0.5), q, S(0.5), S(-0.5));
  tti_grad_ratio<S>(l, q, a, S(0), diag[0][0], diag[0][1]);
  tti_grad_ratio<S>(l, q, S(0), b, diag[kTTIMix - 1][0], diag[kTTIMix - 1][1]);
#pragma unroll
  for (int i = 1; i < kTTIMix - 1; ++i) {
    const S t = S(i) / S(kTTIMix - 1);
    const S v = (S(1) - t) * a - t * b;
    const S u = tti_solve2_above<S>(S(-2) + S(2) * L1 * v + Q2 * v * v, L0 + Q1 * v, Q0, abs_(v));
    tti_grad_ratio<S>(l, q, (u + v) / S(2), (u - v) / S(2), diag[i][0], diag[i][1]);
  }
  return diag[0][0] * diag[kTTIMix - 1][1] - diag[0][1] * diag[kTTIMix - 1][0] > S(0);
}

// D0 = sum of the outer products of T's first D-1 rows, D1 = that of the
// last row (flattened lower triangles); g = [l0, l1, q00, q10, q11, T].
template <class S, int D> __device__ inline void tti_base(const S* g, S* D0, S* D1) {
#pragma clang fp contract(off)
  const S* T = g + 5;
  int k = 0;
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j, ++k) {
      D0[k] = T[i] * T[j];
      D1[k] = T[(D - 1) * D + i] * T[(D - 1) * D + j];
      if constexpr (D == 3) D0[k] += T[D + i] * T[D + j];
    }
}

// Sub-stencil m: Selling decomposition of diag[m][0] D0 + diag[m][1] D1.
template <class S, int D>
__device__ inline void tti_sub(const S* D0, const S* D1, S d0, S d1, S* w, int (&e)[D == 2 ? 3 : 6][3]) {
  constexpr int SD = D * (D + 1) / 2;
  S m[SD];
  {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < SD; ++k) m[k] = d0 * D0[k] + d1 * D1[k];
  }
  if constexpr (D == 2) selling2<S>(m, w, e);
  else selling3<S>(m, w, e);
}

// ------------------------------------------------------------------ TTI factor
// Source factorization of the TTI model (TTI_.h:139-212, GoldenSearch.h,
// EuclideanFactor.h:34-50).  The factor metric is
// fm = [l0, l1, q00, q10, q11, A (D x D, row-major)], the norm N(x) = N0(A x)
// with N0 the primal norm of the dual ball l(s) + 1/2 q(s, s) = 1,
// s = (X^2 [+ Y^2], Z^2).  N0 is approached by a golden-section search over
// the tangent ellipses of that ball: ellipse v (v in [a, -b], the contact
// point x = ((u+v)/2, (u-v)/2) on the curve) has the primal norm
// sqrt(s0 / g0 + s1 / g1), g the gradient ratio at x; the search maximises
// or minimises as the determinant of the first two ellipses says.  Written
// without contraction, as the oracle.
constexpr int kGoldenSteps = 8;  // niter_golden_search (TTI_.h:141-143)

template <class S> struct TTIFactorData {
  S l[2], q[3];
  S L0, L1, Q0, Q1, Q2;  // l and q in the basis (1/2, 1/2), (1/2, -1/2)
  S a, b;                // (a, 0) and (0, b) lie on the curve
};

template <class S> __device__ inline void tti_factor_data(const S* fm, TTIFactorData<S>& t) {
#pragma clang fp contract(off)
  t.l[0] = fm[0];
  t.l[1] = fm[1];
  t.q[0] = fm[2];
  t.q[1] = fm[3];
  t.q[2] = fm[4];
  t.a = tti_solve2_above<S>(S(-2), t.l[0], t.q[0], S(0));
  t.b = tti_solve2_above<S>(S(-2), t.l[1], t.q[2], S(0));
  t.L0 = t.l[0] * S(0.5) + t.l[1] * S(0.5);
  t.L1 = t.l[0] * S(0.5) + t.l[1] * S(-0.5);
  t.Q0 = tti_vqv<S>(S(0.5), S(0.5), t.q, S(0.5), S(0.5));
  t.Q1 = tti_vqv<S>(S(0.5), S(0.5), t.q, S(0.5), S(-0.5));
  t.Q2 = tti_vqv<S>(S(0.5), S(-0.5), t.q, S(0.5), S(-0.5));
}

// Primal norm of the ellipse of parameter v at s; (d0, d1) = 1 / g.
template <class S>
__device__ inline S tti_ellipse_norm(const TTIFactorData<S>& t, S s0, S s1, S v, S& d0, S& d1) {
#pragma clang fp contract(off)
  const S u = tti_solve2_above<S>(S(-2) + S(2) * t.L1 * v + t.Q2 * v * v, t.L0 + t.Q1 * v, t.Q0, abs_(v));
  S g0, g1;
  tti_grad_ratio<S>(t.l, t.q, (u + v) / S(2), (u - v) / S(2), g0, g1);
  d0 = S(1) / g0;
  d1 = S(1) / g1;
  return sqrt_(s0 * d0 + s1 * d1);
}

// N(x); grad (nullable) receives A^T (diag(d) A x) / N(x), the gradient of
// the last ellipse of the search.
template <class S, int D>
__device__ inline S tti_factor_norm(const S* fm, const TTIFactorData<S>& t, const S* x, S* grad) {
#pragma clang fp contract(off)
  const S* A = fm + 5;
  S Ax[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S r = S(0);
#pragma unroll
    for (int j = 0; j < D; ++j) r += A[i * D + j] * x[j];
    Ax[i] = r;
  }
  S s0 = Ax[0] * Ax[0];
  const S s1 = Ax[D - 1] * Ax[D - 1];
  if constexpr (D == 3) s0 += Ax[1] * Ax[1];
  const S phi = S(0.6180339887498949), psi = S(0.3819660112501051);
  S mid[2] = {phi * t.a + psi * (-t.b), psi * t.a + phi * (-t.b)};
  S val[2], e0[2], e1[2];
  val[0] = tti_ellipse_norm<S>(t, s0, s1, mid[0], e0[0], e0[1]);
  val[1] = tti_ellipse_norm<S>(t, s0, s1, mid[1], e1[0], e1[1]);
  const bool mix_min = e0[0] * e1[1] - e0[1] * e1[0] <= S(0);
  S d0 = S(0), d1 = S(0);
  S norm = val[1];
  for (int i = 0; i < kGoldenSteps; ++i) {
    const S dx = mid[1] - mid[0];
    if (mix_min == (val[0] < val[1])) {
      mid[1] = mid[0];
      val[1] = val[0];
      mid[0] -= dx * phi;
      val[0] = tti_ellipse_norm<S>(t, s0, s1, mid[0], d0, d1);
      norm = val[0];
    } else {
      mid[0] = mid[1];
      val[0] = val[1];
      mid[1] += dx * phi;
      val[1] = tti_ellipse_norm<S>(t, s0, s1, mid[1], d0, d1);
      norm = val[1];
    }
  }
  if (grad) {
#pragma unroll
    for (int i = 0; i < D; ++i) Ax[i] = Ax[i] * (((D == 2 ? i : i - 1) <= 0 ? d0 : d1) / norm);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      S r = S(0);
#pragma unroll
      for (int i = 0; i < D; ++i) r += A[i * D + j] * Ax[i];
      grad[j] = r;
    }
  }
  return norm;
}

// Shifts of the neighbours x - e and x + e (fact), and of x - 2e and x + 2e
// (fact2, nullable) of the cell at x (relative to the origin), given
// nx = N(x) and its gradient (generic_factor_sym, EuclideanFactor.h:34-50).
template <class S, int D>
__device__ inline void tti_factor_pair(const S* fm, const TTIFactorData<S>& t, const S* x, S nx,
                                       const S* grad, const int* e, S* fact, S* fact2) {
#pragma clang fp contract(off)
  S xp[D], xm[D];
  S ge = S(0);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    xp[c] = x[c] + S(e[c]);
    xm[c] = x[c] - S(e[c]);
    ge += grad[c] * S(e[c]);
  }
  const S np = tti_factor_norm<S, D>(fm, t, xp, nullptr);
  const S nm = tti_factor_norm<S, D>(fm, t, xm, nullptr);
  fact[0] = -ge + nx - nm;
  fact[1] = ge + nx - np;
  if (fact2) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      xp[c] = xp[c] + S(e[c]);
      xm[c] = xm[c] - S(e[c]);
    }
    const S np2 = tti_factor_norm<S, D>(fm, t, xp, nullptr);
    const S nm2 = tti_factor_norm<S, D>(fm, t, xm, nullptr);
    fact2[0] = S(2) * fact[0] - (nx - S(2) * nm + nm2);
    fact2[1] = S(2) * fact[1] - (nx - S(2) * np + np2);
  }
}

// ------------------------------------------------------------------ curvature models
// The curvature-penalised models on the (x, y, theta) grid (ReedsShepp2.h,
// ReedsSheppForward2.h, Elastica2.h, Dubins2.h with Decomp_v_.h and
// Geometry_.h:199-210).  g = [1/xi, kappa, cos theta, sin theta] per cell in
// grid units; v = (cos theta, sin theta, kappa) is the direction of motion.
// relax is the relaxation of the self outer products (eps^2, eps = 0.1 by
// default).  Written without contraction, as the oracle.

// m = (1 - relax) v v^T + relax |v|^2 Id, lower triangle.
template <class S> __device__ inline void outer_relax3(const S* v, S relax, S* m) {
#pragma clang fp contract(off)
  const S rv2 = relax * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const S rm1 = S(1) - relax;
  int k = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j, ++k) {
      m[k] = v[i] * v[j] * rm1;
      if (i == j) m[k] += rv2;
    }
}

// Decomp_v_.h:7-25: Selling decomposition of the relaxed v v^T; offsets
// deviating from v by more than acos(sqrt(2/3)) get weight 0, the others
// point against v (<v, e> <= 0: the forward neighbour x + e lies upwind).
template <class S> __device__ inline void decomp_v(const S* v, S relax, S* w, int (*e)[3]) {
#pragma clang fp contract(off)
  S m[6];
  outer_relax3<S>(v, relax, m);
  int eb[6][3];
  selling3<S>(m, w, eb);
  const S vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const S cosmin2 = S(2.0 / 3.0);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const S ve = v[0] * S(eb[k][0]) + v[1] * S(eb[k][1]) + v[2] * S(eb[k][2]);
    const S ee = S(eb[k][0] * eb[k][0] + eb[k][1] * eb[k][1] + eb[k][2] * eb[k][2]);
    const bool off = ve * ve < vv * ee * cosmin2;
    const int sg = (!off && ve > S(0)) ? -1 : 1;
    if (off) w[k] = S(0);
#pragma unroll
    for (int c = 0; c < 3; ++c) e[k][c] = sg * eb[k][c];
  }
}

// Fejer's first quadrature rule on [0, pi] with the weight sin(phi), n nodes
// phi_l = pi (l + 1/2) / n: w_l = (2/n) (1 - 2 sum_{j=1}^{n/2} cos(2 j phi_l) /
// (4 j^2 - 1)) (Elastica2.h:13-45 uses the rule with n = 5).  Rounded to the
// precision of the reference's tables, so that the stencils agree with it:
// cosines and sines to 9 decimals (cos(pi/2) is then exactly 0), weights to 6.
template <class S> __device__ inline void fejer_node(int l, int n, S& c, S& s, S& w) {
  const double phi = 3.141592653589793 * (double(l) + 0.5) / double(n);
  double acc = 0.;
  for (int j = 1; j <= n / 2; ++j) acc += cos(2. * j * phi) / double(4 * j * j - 1);
  c = S(rint(cos(phi) * 1e9) / 1e9);
  s = S(rint(sin(phi) * 1e9) / 1e9);
  w = S(rint(2. / n * (1. - 2. * acc) * 1e6) / 1e6);
}
constexpr int kNFejer = 5;  // nFejer_macro of Elastica2 (kernel_traits.py:30-33)

// Stencil of a curvature model: NMIX sub-schemes of NSYM offset slots each
// (slot k of mix m at m * NSYM + k; the last NFWD slots are forward
// offsets).  Returns mix_is_min (Dubins2: the smaller of its two updates).
// convex: the Convex variants, whose vehicle turns left only (Dubins2: the
// right turn becomes straight motion; Elastica2: the nodes of right turns
// weigh 0, the straight one half; ReedsSheppForward2: the rotation slot is
// forward, which the neighbour codes express).
template <class S, int MODEL>
__device__ inline bool curv_stencil(const S* g, S relax, bool convex, S* w, int (*e)[3]) {
#pragma clang fp contract(off)
  const S ixi = g[0], kappa = g[1], cT = g[2], sT = g[3];
  if constexpr (MODEL == AGD_REEDSSHEPP2) {
    // ReedsShepp2.h:13-36
    const S v[3] = {cT, sT, kappa};
    S m[6];
    outer_relax3<S>(v, relax, m);
    const S m5 = v[2] * v[2] + ixi * ixi;
    m[5] = m[5] > m5 ? m[5] : m5;
    int eb[6][3];
    selling3<S>(m, w, eb);
    const S wv[3] = {v[1], -v[0], S(0)};
    const S ww = wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2];
    const S cut = S(1) - S(2.0 / 3.0);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const S we = wv[0] * S(eb[k][0]) + wv[1] * S(eb[k][1]) + wv[2] * S(eb[k][2]);
      const S ee = S(eb[k][0] * eb[k][0] + eb[k][1] * eb[k][1] + eb[k][2] * eb[k][2]);
      if (we * we >= ee * ww * cut) w[k] = S(0);
#pragma unroll
      for (int c = 0; c < 3; ++c) e[k][c] = eb[k][c];
    }
    return true;
  } else if constexpr (MODEL == AGD_REEDSSHEPPFWD2) {
    // ReedsSheppForward2.h:19-33: pure rotation (symmetric), then forward motion
    w[0] = ixi * ixi;
    e[0][0] = 0;
    e[0][1] = 0;
    e[0][2] = -1;
    const S v[3] = {cT, sT, kappa};
    decomp_v<S>(v, relax, w + 1, e + 1);
    return true;
  } else if constexpr (MODEL == AGD_DUBINS2) {
    // Dubins2.h:17-36: left and right turns at the largest curvature
    const S vl[3] = {cT, sT, kappa + ixi};
    const S vr[3] = {cT, sT, convex ? kappa : kappa - ixi};
    decomp_v<S>(vl, relax, w, e);
    decomp_v<S>(vr, relax, w + 6, e + 6);
    return true;
  } else {
    // Elastica2.h:51-79: the directions (sin phi v_xy, sin phi kappa + cos phi / xi)
#pragma unroll 1
    for (int l = 0; l < kNFejer; ++l) {
      S cP, sP, wF;
      fejer_node<S>(l, kNFejer, cP, sP, wF);
      if (convex) wF *= 2 * l < kNFejer - 1 ? S(1) : (2 * l == kNFejer - 1 ? S(0.5) : S(0));
      const S v[3] = {sP * cT, sP * sT, sP * kappa + cP * ixi};
      decomp_v<S>(v, relax, w + 6 * l, e + 6 * l);
#pragma unroll
      for (int i = 0; i < 6; ++i) w[6 * l + i] *= wF;
    }
    return true;
  }
}

// ------------------------------------------------------------------ update
// Compare-exchange of the sorting network (NetworkSort.h:17-22): position i
// keeps its pair when v[i] < v[j], otherwise the pairs swap (ties swap).
template <class S>
__device__ inline void cex(S& vi, S& wi, S& vj, S& wj) {
  const bool keep = vi < vj;
  const S a = keep ? vi : vj, b = keep ? vj : vi;
  const S c = keep ? wi : wj, d = keep ? wj : wi;
  vi = a; vj = b; wi = c; wj = d;
}

// Batcher's odd-even merge sort network on the power of two P >= N, without
// the comparators that touch a padding position (index >= N, holding +inf,
// which every comparator would leave in place): 7 offsets take 16 comparators
// instead of 21 for odd-even transposition, 30 offsets 178 instead of 435.
constexpr int next_pow2(int n) { return n <= 1 ? 1 : 2 * next_pow2((n + 1) / 2); }
template <int N> struct MergeNet {
  static constexpr int P = next_pow2(N);
  int n;
  int a[P * 8], b[P * 8];
  constexpr MergeNet() : n(0), a{}, b{} {
    for (int p = 1; p < P; p <<= 1)
      for (int k = p; k >= 1; k >>= 1)
        for (int j = k % p; j <= P - 1 - k; j += 2 * k)
          for (int i = 0; i <= (k - 1 < P - j - k - 1 ? k - 1 : P - j - k - 1); ++i)
            if ((i + j) / (2 * p) == (i + j + k) / (2 * p) && i + j + k < N) {
              a[n] = i + j;
              b[n] = i + j + k;
              ++n;
            }
  }
};

template <class S, int N> __device__ inline void sort_pairs(S (&v)[N], S (&w)[N]) {
  if constexpr (N == 1) {
    // nothing to sort
  } else if constexpr (N == 2) {
    cex(v[0], w[0], v[1], w[1]);
  } else if constexpr (N == 3) {
    cex(v[1], w[1], v[2], w[2]); cex(v[0], w[0], v[2], w[2]); cex(v[0], w[0], v[1], w[1]);
  } else if constexpr (N == 6) {
    cex(v[1], w[1], v[2], w[2]); cex(v[4], w[4], v[5], w[5]);
    cex(v[0], w[0], v[2], w[2]); cex(v[3], w[3], v[5], w[5]);
    cex(v[0], w[0], v[1], w[1]); cex(v[3], w[3], v[4], w[4]); cex(v[2], w[2], v[5], w[5]);
    cex(v[0], w[0], v[3], w[3]); cex(v[1], w[1], v[4], w[4]);
    cex(v[2], w[2], v[4], w[4]); cex(v[1], w[1], v[3], w[3]);
    cex(v[2], w[2], v[3], w[3]);
  } else {
    // other lengths (the curvature models' 7 and 30)
    constexpr MergeNet<N> net{};
    constexpr int count = net.n;
#pragma unroll
    for (int c = 0; c < count; ++c) cex(v[net.a[c]], w[net.a[c]], v[net.b[c]], w[net.b[c]]);
  }
}

// HFM.h:165-195: smallest solution of sum_k w_k (u - v_k)_+^2 = rhs^2 over
// the sorted neighbour values, growing the active set while the quadratic is
// non-positive at the next value.  Inputs by value (sorted copies).
template <class S, int N>
__device__ inline S hfm_solve(S (&v)[N], S (&w)[N], S rhs) {
  sort_pairs<S, N>(v, w);
  const S vmin = v[0];
  S a = w[0], b = S(0), c = -rhs * rhs;
  bool go = true;
#pragma unroll
  for (int k = 1; k < N; ++k) {
    const S t = v[k] - vmin;
    go = go && !((a * t - S(2) * b) * t + c > S(0));
    if (go) {
      a += w[k];
      b += w[k] * t;
      c += w[k] * t * t;
    }
  }
  const S delta = b * b - a * c;
  const S sdelta = sqrt_fast(delta > S(0) ? delta : S(0));
  const S value = (a == S(0)) ? inf_v<S>() : div_fast(b + sdelta, a);
  return (vmin == inf_v<S>()) ? vmin : vmin + value;
}

// ------------------------------------------------------------------ walls
__device__ inline int div_round_closest(int n, int d) {
  return ((n < 0) ^ (d < 0)) ? ((n - d / 2) / d) : ((n + d / 2) / d);
}

// Walls.h:18-50.  wx = wall distance at x (255 = far).  `lookup(w)` returns
// the wall distance at x + w (255 outside the domain).
template <int D, class Lookup>
__device__ inline bool visible(const int* e, int wx, Lookup lookup) {
  if (wx == 255) return true;
  int l1 = 0, linf = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const int a = e[k] < 0 ? -e[k] : e[k];
    l1 += a;
    linf = a > linf ? a : linf;
  }
  if (l1 < wx) return true;
  for (int t = 1; t <= linf; ++t) {
    int w[3] = {0, 0, 0};
    int diffl1 = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      w[k] = div_round_closest(e[k] * t, linf);
      const int r = e[k] - w[k];
      diffl1 += r < 0 ? -r : r;
    }
    const int value = lookup(w);
    if (diffl1 < value) return true;
    if (value == 0) return false;
  }
  return true;
}

}  // namespace agd
