// agd_oracle.cpp -- CPU restatement of the reference eikonal scheme.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle against which the HIP
// product path is checked.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product package never does.
//
// What it restates (reference paths relative to the AGD source tree):
//   * Selling decomposition, 2D/3D, CUDA ordering:
//       agd/Eikonal/HFM_CUDA/cuda/Geometry2.h:21-48, Geometry3.h:23-58,
//       scal_vmv summation order Geometry_.h:144-156, canonical superbase
//       Geometry_.h:237-243.
//   * Scheme per model: Isotropic_.h:20-41 (isotropic / diagonal),
//       Riemann_.h:16-29 (Riemann, Rander = Riemann + drift),
//       drift shift FiniteDifferences.h:41-45 (fact[0]+=s, fact[1]-=s),
//       TTI_.h:41-134 (7 tangent ellipses of the dual unit ball, one
//       Selling decomposition each, min or max of the 7 updates as
//       HFMIter.h:29-56 with mix_is_min from diags()).
//   * Neighbour values (min over -e/+e, ties keep -e), network sort and the
//       upwind quadratic update: HFM.h:28-207, NetworkSort.h:17-60.
//   * Convergence test |u_old-u_new| <= atol + rtol|u_new|: Update.h:211-219.
//   * Flow weights w_k (u - v_k)_+ and flow vector: HFM.h:197-207,
//       Update.h:171-209.
//   * Reverse AD system (D - W)^T lambda = s: _PostProcess.py:155-219,
//       Kernel_LinearUpdate.h:99-216 (solved here exactly, in causal order).
//   * Walls: visibility walk Walls.h:18-50 and L1 wall distance
//       _SetGeometry.py:262-282 / Kernel_InfConvolution.h:66-91.
//   * Curvature-penalised models on (x, y, theta) grids: ReedsShepp2.h:13-36,
//       ReedsSheppForward2.h:19-33, Elastica2.h:51-79 (Fejer rule, 5 nodes),
//       Dubins2.h:17-36, with decomp_v (Decomp_v_.h:7-21), the relaxed self
//       outer product (Geometry_.h:199-210) and forward offsets, whose value
//       is u(x + e) alone (HFM.h:64-75, FiniteDifferences.h:48).
//
// Solvers:
//   * orc_solve_fm     : Dijkstra-like fast marching in FP64 -- the exact
//                        fixed point of the causal upwind scheme (restated
//                        from Notebooks_NonDiv/EikonalEulerian.ipynb cell 35).
//   * orc_solve_jacobi : strict global Jacobi iteration (FP32 or FP64) with the
//                        reference stopping rule; used to show that the fixed
//                        point does not depend on the schedule.
//
// Arithmetic conventions: scal_vmv accumulates with explicit fused
// multiply-add (the contracted form that the reference's JIT compiler emits
// with its default fmad=true); the product kernels do the same, so integer
// offsets agree bit for bit.  Arrays are cell-major, row-major (last axis
// fastest).  Geometry is component-major: geom[k*N + n].

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

template <class S> static inline S Inf() { return std::numeric_limits<S>::infinity(); }
static inline float fmaT(float a, float b, float c) { return std::fmaf(a, b, c); }
static inline double fmaT(double a, double b, double c) { return std::fma(a, b, c); }

// ---------------------------------------------------------------- Selling
// scal_vmv(x, m, y) = sum_{i>=j} (i==j ? x_i y_i : x_i y_j + x_j y_i) m_ij,
// integer products first, then one fma per matrix coefficient
// (Geometry_.h:144-156).
template <int D, class S>
static S scal_vmv(const int* x, const S* m, const int* y) {
  S r = 0;
  int k = 0;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j <= i; ++j, ++k) {
      const int p = (i == j) ? x[i] * y[i] : (x[i] * y[j] + x[j] * y[i]);
      r = fmaT(S(p), m[k], r);
    }
  return r;
}

static const int kSellingMaxIter = 100;  // Geometry2.h:13 / Geometry3.h:10

// Geometry2.h:21-36: cyclic (i,j,k); if <b_i,D b_j> > 0 then b_k = b_i - b_j,
// b_i = -b_i; stop after 3 consecutive non-reducing steps.
template <class S>
static void obtuse_superbase2(const S* m, int sb[3][2]) {
  sb[0][0] = 1; sb[0][1] = 0;
  sb[1][0] = 0; sb[1][1] = 1;
  sb[2][0] = -1; sb[2][1] = -1;
  for (int iter = 0, red = 0; iter < kSellingMaxIter && red < 3; ++iter, ++red) {
    const int i = iter % 3, j = (iter + 1) % 3, k = (iter + 2) % 3;
    if (scal_vmv<2, S>(sb[i], m, sb[j]) > 0) {
      for (int l = 0; l < 2; ++l) sb[k][l] = sb[i][l] - sb[j][l];
      for (int l = 0; l < 2; ++l) sb[i][l] = -sb[i][l];
      red = 0;
    }
  }
}

// Geometry2.h:40-48: w_r = max(0, -<b_i,D b_j>), e_r = perp(b_k), perp(x)=(-x1,x0).
template <class S>
static void selling2(const S* m, S* w, int e[][3]) {
  int sb[3][2];
  obtuse_superbase2<S>(m, sb);
  for (int r = 0; r < 3; ++r) {
    const int i = r, j = (r + 1) % 3, k = (r + 2) % 3;
    const S s = -scal_vmv<2, S>(sb[i], m, sb[j]);
    w[r] = s > S(0) ? s : S(0);
    e[r][0] = -sb[k][1];
    e[r][1] = sb[k][0];
    e[r][2] = 0;
  }
}

// Geometry3.h:19-20: pairs (i,j) range over all 6 pairs, (k,l) the complement.
static const int kPerm3[6][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {0, 3, 1, 2},
                                 {1, 2, 0, 3}, {1, 3, 0, 2}, {2, 3, 0, 1}};

// Geometry3.h:23-39: if <b_i,D b_j> > 0 then b_k += b_i, b_l += b_i, b_i = -b_i.
template <class S>
static void obtuse_superbase3(const S* m, int sb[4][3]) {
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 3; ++b) sb[a][b] = (a == 3) ? -1 : (a == b ? 1 : 0);
  for (int iter = 0, red = 0; iter < kSellingMaxIter && red < 6; ++iter, ++red) {
    const int* p = kPerm3[iter % 6];
    const int i = p[0], j = p[1];
    if (scal_vmv<3, S>(sb[i], m, sb[j]) > 0) {
      const int k = p[2], l = p[3];
      for (int c = 0; c < 3; ++c) {
        sb[k][c] += sb[i][c];
        sb[l][c] += sb[i][c];
        sb[i][c] = -sb[i][c];
      }
      red = 0;
    }
  }
}

// Geometry3.h:43-58: w_r = max(0,-<b_i,D b_j>), e_r = b_k x b_l.
template <class S>
static void selling3(const S* m, S* w, int e[][3]) {
  int sb[4][3];
  obtuse_superbase3<S>(m, sb);
  for (int r = 0; r < 6; ++r) {
    const int* p = kPerm3[r];
    const int i = p[0], j = p[1], k = p[2], l = p[3];
    const S s = -scal_vmv<3, S>(sb[i], m, sb[j]);
    w[r] = s > S(0) ? s : S(0);
    for (int c = 0; c < 3; ++c) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      e[r][c] = sb[k][c1] * sb[l][c2] - sb[k][c2] * sb[l][c1];
    }
  }
}

// ---------------------------------------------------------------- models
enum Model {
  ISOTROPIC = 0, DIAGONAL = 1, RIEMANN = 2, RANDER = 3, TTI = 4,
  REEDSSHEPP2 = 5, REEDSSHEPPFWD2 = 6, ELASTICA2 = 7, DUBINS2 = 8
};
static const int kNMix = 7;  // TTI: nmix_macro (kernel_traits.py:47-52)
static const int kNFejer = 5;  // Elastica2: nFejer_macro (kernel_traits.py:30-33)
static const int kMaxAct = 30;  // offsets of one sub-stencil (Elastica2: 5 x 6)
static const int kMaxSlots = 42;  // all sub-stencils (TTI: 7 x 6)

static int symdim(int d) { return d * (d + 1) / 2; }
static bool curvature(int model) { return model >= REEDSSHEPP2 && model <= DUBINS2; }
// Offsets of one sub-stencil, each a pair x -+ e; the last nfwd_of of them
// are forward offsets (x + e only).
static int nsym_of(int model, int d) {
  switch (model) {
    case ISOTROPIC:
    case DIAGONAL: return d;
    case REEDSSHEPP2: return 6;
    case REEDSSHEPPFWD2: return 7;
    case ELASTICA2: return 6 * kNFejer;
    case DUBINS2: return 6;
  }
  return symdim(d);
}
static int nfwd_of(int model) {
  switch (model) {
    case REEDSSHEPPFWD2: return 6;
    case ELASTICA2: return 6 * kNFejer;
    case DUBINS2: return 6;
  }
  return 0;
}
static int ngeom_of(int model, int d) {
  switch (model) {
    case ISOTROPIC: return 0;
    case DIAGONAL: return d;
    case RIEMANN: return symdim(d);
    case RANDER: return symdim(d) + d;
    case TTI: return 5 + d * d;
  }
  return curvature(model) ? 4 : 0;
}
static int nmix_of(int model) { return model == TTI ? kNMix : (model == DUBINS2 ? 2 : 1); }

// nmix sub-stencils of nsym weights/offsets each (sub-stencil m at
// [m*nsym, (m+1)*nsym)); the update is the min (mix_is_min) or the max of the
// sub-stencils' updates.
template <class S>
struct Stencil {
  int nsym, nmix, nfwd;
  bool mix_is_min;
  S w[kMaxSlots];
  int e[kMaxSlots][3];
  S drift[3];
  bool has_drift;
};

// ---------------------------------------------------------------- TTI
// TTI_.h:41-49: both roots of a + 2 b t + c t^2 (non-negative discriminant).
template <class S>
static void solve2(S a, S b, S c, S r[2]) {
  const S sdelta = std::sqrt(b * b - a * c);
  const S u = -b + sdelta, v = -sdelta - b;
  if (std::fabs(c) > std::fabs(a)) { r[0] = u / c; r[1] = v / c; return; }
  if (a != S(0)) { r[0] = a / u; r[1] = a / v; return; }
  r[0] = 0; r[1] = Inf<S>();
}

// TTI_.h:53-58: smallest root above the threshold.
template <class S>
static S solve2_above(S a, S b, S c, S above) {
  S r[2];
  solve2<S>(a, b, c, r);
  const bool ordered = r[0] < r[1];
  const S rmin = ordered ? r[0] : r[1], rmax = ordered ? r[1] : r[0];
  return rmin >= above ? rmin : rmax;
}

// TTI_.h:67-72: g = df(x) / <x, df(x)> with f(x) = C + 2<l,x> + <qx,x>,
// q = (q00, q10, q11).
template <class S>
static void grad_ratio(const S* l, const S* q, const S* x, S* g) {
  S h0 = q[0] * x[0] + q[1] * x[1];
  S h1 = q[1] * x[0] + q[2] * x[1];
  h0 = l[0] + h0;
  h1 = l[1] + h1;
  const S xhg = x[0] * h0 + x[1] * h1;
  g[0] = h0 / xhg;
  g[1] = h1 / xhg;
}

// x^T q y for a symmetric 2x2 q = (q00, q10, q11) (Geometry_.h scal_vmv order).
template <class S>
static S scal_vmv2f(const S* x, const S* q, const S* y) {
  return x[0] * q[0] * y[0] + (x[0] * y[1] + x[1] * y[0]) * q[1] + x[1] * q[2] * y[1];
}

// TTI_.h:76-107: samples the curve f(x) = 0, x >= 0 at nmix points from
// (a,0) to (0,b) and returns diag_s[i] = grad f / <x, grad f> there;
// mix_is_min = det(diag_s[0], diag_s[nmix-1]) > 0.
template <class S>
static bool tti_diags(const S* l, const S* q, S diag_s[kNMix][2]) {
  const S a = solve2_above<S>(S(-2), l[0], q[0], S(0));
  const S b = solve2_above<S>(S(-2), l[1], q[2], S(0));
  const S e0[2] = {S(0.5), S(0.5)}, e1[2] = {S(0.5), S(-0.5)};
  const S L0 = l[0] * e0[0] + l[1] * e0[1], L1 = l[0] * e1[0] + l[1] * e1[1];
  const S Q0 = scal_vmv2f<S>(e0, q, e0), Q1 = scal_vmv2f<S>(e0, q, e1), Q2 = scal_vmv2f<S>(e1, q, e1);
  S x_s[kNMix][2];
  x_s[0][0] = a; x_s[0][1] = 0;
  x_s[kNMix - 1][0] = 0; x_s[kNMix - 1][1] = b;
  for (int i = 1; i < kNMix - 1; ++i) {
    const S t = S(i) / S(kNMix - 1);
    const S v = (1 - t) * a - t * b;
    const S u = solve2_above<S>(S(-2) + S(2) * L1 * v + Q2 * v * v, L0 + Q1 * v, Q0, std::fabs(v));
    x_s[i][0] = (u + v) / 2;
    x_s[i][1] = (u - v) / 2;
  }
  for (int i = 0; i < kNMix; ++i) grad_ratio<S>(l, q, x_s[i], diag_s[i]);
  return diag_s[0][0] * diag_s[kNMix - 1][1] - diag_s[0][1] * diag_s[kNMix - 1][0] > 0;
}

// TTI_.h:112-134: geometry [l0, l1, q00, q10, q11, T (d x d, row-major)];
// D0 = sum of the outer products of T's first d-1 rows, D1 = that of its
// last row, sub-stencil i = Selling(diag_i[0] D0 + diag_i[1] D1).
template <class S>
static void tti_scheme(int d, const S* g, Stencil<S>& st) {
  const S* l = g;
  const S* q = g + 2;
  const S* T = g + 5;
  S diag_s[kNMix][2];
  st.mix_is_min = tti_diags<S>(l, q, diag_s);
  const int sd = symdim(d);
  S D0[6], D1[6];
  for (int i = 0, k = 0; i < d; ++i)
    for (int j = 0; j <= i; ++j, ++k) {
      D0[k] = T[i] * T[j];
      D1[k] = T[(d - 1) * d + i] * T[(d - 1) * d + j];
      if (d == 3) D0[k] += T[d + i] * T[d + j];
    }
  for (int m = 0; m < kNMix; ++m) {
    S D[6];
    for (int k = 0; k < sd; ++k) D[k] = diag_s[m][0] * D0[k] + diag_s[m][1] * D1[k];
    if (d == 2) selling2<S>(D, st.w + m * sd, st.e + m * sd);
    else selling3<S>(D, st.w + m * sd, st.e + m * sd);
  }
}

// ---------------------------------------------------------------- TTI factor
// TTI_.h:139-212 with GoldenSearch.h and EuclideanFactor.h:34-50: the source
// factorization of the TTI model uses the norm N(x) = N0(A x) of the factor
// metric fm = [l0, l1, q00, q10, q11, A (d x d, row-major)], N0 approached by
// niter_golden_search = 8 golden-section steps over the tangent ellipses of
// the dual ball.  Ellipse v, v in [a, -b], touches the curve at
// x = ((u+v)/2, (u-v)/2) (u as in tti_diags) and has the primal norm
// sqrt(s0 / g0 + s1 / g1) with s = (X^2 [+ Y^2], Z^2) of A x and g the
// gradient ratio at x.  The search maximises, or minimises when the first
// two ellipses' reciprocal ratios have a non-positive determinant.
static const int kGoldenSteps = 8;

template <class S>
struct TTIFactorData {
  S l[2], q[3], L[2], Q[3], a, b;
};

// TTI_.h:86-97 (tti_data_init).
template <class S>
static void tti_factor_data(const S* fm, TTIFactorData<S>& t) {
  for (int i = 0; i < 2; ++i) t.l[i] = fm[i];
  for (int i = 0; i < 3; ++i) t.q[i] = fm[2 + i];
  t.a = solve2_above<S>(S(-2), t.l[0], t.q[0], S(0));
  t.b = solve2_above<S>(S(-2), t.l[1], t.q[2], S(0));
  const S e0[2] = {S(0.5), S(0.5)}, e1[2] = {S(0.5), S(-0.5)};
  t.L[0] = t.l[0] * e0[0] + t.l[1] * e0[1];
  t.L[1] = t.l[0] * e1[0] + t.l[1] * e1[1];
  t.Q[0] = scal_vmv2f<S>(e0, t.q, e0);
  t.Q[1] = scal_vmv2f<S>(e0, t.q, e1);
  t.Q[2] = scal_vmv2f<S>(e1, t.q, e1);
}

// TTI_.h:148-156: the ellipse of parameter v; diag = 1 / g.
template <class S>
static S tti_ellipse(const TTIFactorData<S>& t, const S s[2], S v, S diag[2]) {
  const S u = solve2_above<S>(S(-2) + S(2) * t.L[1] * v + t.Q[2] * v * v, t.L[0] + t.Q[1] * v, t.Q[0],
                              std::fabs(v));
  const S x[2] = {(u + v) / 2, (u - v) / 2};
  S g[2];
  grad_ratio<S>(t.l, t.q, x, g);
  diag[0] = S(1) / g[0];
  diag[1] = S(1) / g[1];
  return std::sqrt(s[0] * diag[0] + s[1] * diag[1]);
}

// TTI_.h:159-189: N(x) by the golden-section search, and (grad non-null) the
// gradient A^T diag(d) A x / N(x) of the last ellipse evaluated.
template <class S>
static S tti_factor_norm(int d, const S* fm, const TTIFactorData<S>& t, const S* x, S* grad) {
  const S* A = fm + 5;
  S Ax[3] = {0, 0, 0};
  for (int i = 0; i < d; ++i) {
    S r = 0;
    for (int j = 0; j < d; ++j) r += A[i * d + j] * x[j];
    Ax[i] = r;
  }
  S s[2] = {Ax[0] * Ax[0], Ax[d - 1] * Ax[d - 1]};
  if (d == 3) s[0] += Ax[1] * Ax[1];
  const S phi = S(0.6180339887498949), psi = S(0.3819660112501051);
  const S bounds[2] = {t.a, -t.b};
  S mid[2] = {phi * bounds[0] + psi * bounds[1], psi * bounds[0] + phi * bounds[1]};
  S dg[2][2], diag[2] = {0, 0};
  S val[2] = {tti_ellipse<S>(t, s, mid[0], dg[0]), tti_ellipse<S>(t, s, mid[1], dg[1])};
  const bool mix_is_min = dg[0][0] * dg[1][1] - dg[0][1] * dg[1][0] <= 0;
  int next = 1;
  for (int i = 0; i < kGoldenSteps; ++i) {
    const S dx = mid[1] - mid[0];
    if (mix_is_min == (val[0] < val[1])) {
      mid[1] = mid[0];
      val[1] = val[0];
      mid[0] -= dx * phi;
      next = 0;
    } else {
      mid[0] = mid[1];
      val[0] = val[1];
      mid[1] += dx * phi;
      next = 1;
    }
    val[next] = tti_ellipse<S>(t, s, mid[next], diag);
  }
  const S norm = val[next];
  if (grad) {
    for (int i = 0; i < d; ++i) Ax[i] = Ax[i] * (diag[d == 2 ? i : std::max(i - 1, 0)] / norm);
    for (int j = 0; j < d; ++j) {
      S r = 0;
      for (int i = 0; i < d; ++i) r += A[i * d + j] * Ax[i];
      grad[j] = r;
    }
  }
  return norm;
}

// TTI_.h:191-211 and generic_factor_sym (EuclideanFactor.h:34-50): shifts of
// x - e, x + e and (f2) of x - 2e, x + 2e, given nx = N(x) and its gradient.
template <class S>
static void tti_factor_pair(int d, const S* fm, const TTIFactorData<S>& t, const S* x, S nx,
                            const S* grad, const int* e, S f[2], S* f2) {
  S xp[3] = {0, 0, 0}, xm[3] = {0, 0, 0}, ge = 0;
  for (int c = 0; c < d; ++c) {
    xp[c] = x[c] + S(e[c]);
    xm[c] = x[c] - S(e[c]);
    ge += grad[c] * S(e[c]);
  }
  const S np = tti_factor_norm<S>(d, fm, t, xp, nullptr);
  const S nm = tti_factor_norm<S>(d, fm, t, xm, nullptr);
  f[0] = -ge + nx - nm;
  f[1] = ge + nx - np;
  if (f2) {
    for (int c = 0; c < d; ++c) {
      xp[c] = xp[c] + S(e[c]);
      xm[c] = xm[c] - S(e[c]);
    }
    const S np2 = tti_factor_norm<S>(d, fm, t, xp, nullptr);
    const S nm2 = tti_factor_norm<S>(d, fm, t, xm, nullptr);
    f2[0] = 2 * f[0] - (nx - 2 * nm + nm2);
    f2[1] = 2 * f[1] - (nx - 2 * np + np2);
  }
}

// Isotropic_.h:20-41, Riemann_.h:16-29 and TTI_.h:112-134.
// ---------------------------------------------------------------- curvature
// Geometry_.h:199-210: m = (1 - relax) v v^T + relax |v|^2 Id (lower
// triangle, row by row), each product rounded separately.
template <class S>
static void outer_relax(const S* v, S relax, S* m) {
  const S rv2 = relax * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const S rm1 = S(1) - relax;
  int k = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j <= i; ++j, ++k) {
      m[k] = v[i] * v[j] * rm1;
      if (i == j) m[k] += rv2;
    }
}

// Decomp_v_.h:7-21: the Selling decomposition of the relaxed v v^T, with
// weight 0 for offsets e such that <v,e>^2 < (2/3) |v|^2 |e|^2, and the kept
// offsets turned so that <v,e> <= 0.
template <class S>
static void decomp_v(const S* v, S relax, S* w, int (*e)[3]) {
  S m[6];
  outer_relax<S>(v, relax, m);
  selling3<S>(m, w, e);
  const S vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  for (int k = 0; k < 6; ++k) {
    const S ve = v[0] * S(e[k][0]) + v[1] * S(e[k][1]) + v[2] * S(e[k][2]);
    const S ee = S(e[k][0] * e[k][0] + e[k][1] * e[k][1] + e[k][2] * e[k][2]);
    if (ve * ve < vv * ee * S(2.0 / 3.0)) {
      w[k] = 0;
      continue;
    }
    if (ve > 0)
      for (int c = 0; c < 3; ++c) e[k][c] = -e[k][c];
  }
}

// Fejer's first rule for the integral over [0, pi] of f(phi) sin(phi): nodes
// phi_l = pi (l + 1/2) / n, weights (2/n) (1 - 2 sum_{j=1}^{n/2}
// cos(2 j phi_l) / (4 j^2 - 1)).  Rounded as the reference's table
// (Elastica2.h:13-45, n = 5): cosines and sines to 9 decimals (so that
// cos(pi/2) is exactly 0), weights to 6.
static void fejer(int l, int n, double& c, double& s, double& w) {
  const double phi = 3.141592653589793 * (l + 0.5) / n;
  double acc = 0;
  for (int j = 1; j <= n / 2; ++j) acc += std::cos(2. * j * phi) / double(4 * j * j - 1);
  c = std::nearbyint(std::cos(phi) * 1e9) / 1e9;
  s = std::nearbyint(std::sin(phi) * 1e9) / 1e9;
  w = std::nearbyint(2. / n * (1. - 2. * acc) * 1e6) / 1e6;
}

// Scheme of a curvature model at a cell with geometry g = [1/xi, kappa,
// cos theta, sin theta] (grid units); relax = eps^2.  convex: the Convex
// variants (kernel_traits.py:129-130), whose vehicle turns left only:
// ReedsSheppForward2's rotation becomes a forward offset
// (ReedsSheppForward2.h:10-12), Dubins2's right turn straight motion
// (Dubins2.h:29-33), Elastica2's nodes weigh 1, 1/2, 0 for left turns,
// straight motion and right turns (Elastica2.h:69-73).
template <class S>
static void curvature_scheme(int model, const S* g, S relax, bool convex, Stencil<S>& st) {
  const S ixi = g[0], kappa = g[1], cT = g[2], sT = g[3];
  if (model == REEDSSHEPP2) {
    // ReedsShepp2.h:13-36: the angular coefficient raised to kappa^2 + 1/xi^2,
    // offsets nearly orthogonal to the motion (w = v x e_theta) pruned.
    const S v[3] = {cT, sT, kappa};
    S m[6];
    outer_relax<S>(v, relax, m);
    const S m5 = v[2] * v[2] + ixi * ixi;
    if (m5 > m[5]) m[5] = m5;
    selling3<S>(m, st.w, st.e);
    const S wv[3] = {v[1], -v[0], 0};
    const S ww = wv[0] * wv[0] + wv[1] * wv[1];
    for (int k = 0; k < 6; ++k) {
      const int* e = st.e[k];
      const S we = wv[0] * S(e[0]) + wv[1] * S(e[1]);
      const S ee = S(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      if (we * we >= ee * ww * (S(1) - S(2.0 / 3.0))) st.w[k] = 0;
    }
  } else if (model == REEDSSHEPPFWD2) {
    // ReedsSheppForward2.h:19-33: in-place rotation, then forward motion
    st.w[0] = ixi * ixi;
    st.e[0][0] = 0;
    st.e[0][1] = 0;
    st.e[0][2] = -1;
    const S v[3] = {cT, sT, kappa};
    decomp_v<S>(v, relax, st.w + 1, st.e + 1);
  } else if (model == DUBINS2) {
    // Dubins2.h:17-36: turning left or right at curvature 1/xi, the smaller
    // of the two updates
    const S vl[3] = {cT, sT, kappa + ixi}, vr[3] = {cT, sT, convex ? kappa : kappa - ixi};
    decomp_v<S>(vl, relax, st.w, st.e);
    decomp_v<S>(vr, relax, st.w + 6, st.e + 6);
  } else {
    // Elastica2.h:51-79: the directions (sin phi cos theta, sin phi sin
    // theta, sin phi kappa + cos phi / xi) at the Fejer nodes, weights times
    // the node's Fejer weight
    for (int l = 0; l < kNFejer; ++l) {
      double c, sn, wf;
      fejer(l, kNFejer, c, sn, wf);
      const S cP = S(c), sP = S(sn);
      S wF = S(wf);
      if (convex) wF *= 2 * l < kNFejer - 1 ? S(1) : (2 * l == kNFejer - 1 ? S(0.5) : S(0));
      const S v[3] = {sP * cT, sP * sT, sP * kappa + cP * ixi};
      decomp_v<S>(v, relax, st.w + 6 * l, st.e + 6 * l);
      for (int i = 0; i < 6; ++i) st.w[6 * l + i] *= wF;
    }
  }
}

static double g_curvature_relax = 0.01;  // eps^2 (eps = 0.1, _Kernel.py:319-321)
static bool g_curvature_convex = false;  // the Convex variants

template <class S>
static void make_stencil(int model, int d, const S* g, const S* iso_w, Stencil<S>& st) {
  st.has_drift = false;
  st.nsym = nsym_of(model, d);
  st.nmix = nmix_of(model);
  st.nfwd = nfwd_of(model);
  st.mix_is_min = true;
  for (int k = 0; k < kMaxSlots; ++k) {
    st.w[k] = 0;
    st.e[k][0] = st.e[k][1] = st.e[k][2] = 0;
  }
  st.drift[0] = st.drift[1] = st.drift[2] = 0;
  if (curvature(model)) {
    curvature_scheme<S>(model, g, S(g_curvature_relax), g_curvature_convex, st);
    if (g_curvature_convex && model == REEDSSHEPPFWD2) st.nfwd = st.nsym;  // the rotation too
    return;
  }
  if (model == TTI) {
    st.drift[0] = st.drift[1] = st.drift[2] = 0;
    tti_scheme<S>(d, g, st);
    return;
  }
  st.drift[0] = st.drift[1] = st.drift[2] = 0;
  if (model == ISOTROPIC || model == DIAGONAL) {
    for (int k = 0; k < d; ++k) {
      st.w[k] = (model == ISOTROPIC) ? iso_w[k] : g[k];
      st.e[k][k] = 1;
    }
    return;
  }
  if (d == 2) selling2<S>(g, st.w, st.e);
  else selling3<S>(g, st.w, st.e);
  if (model == RANDER) {
    st.has_drift = true;
    for (int k = 0; k < d; ++k) st.drift[k] = g[symdim(d) + k];
  }
}

// ---------------------------------------------------------------- sorting
// Sorting networks of NetworkSort.h:17-60 (n = 2, 3, 6); the comparison keeps
// order[i] when values[order[i]] < values[order[j]] and swaps otherwise.
static const int kNet2[][2] = {{0, 1}};
static const int kNet3[][2] = {{1, 2}, {0, 2}, {0, 1}};
static const int kNet6[][2] = {{1, 2}, {4, 5}, {0, 2}, {3, 5}, {0, 1}, {3, 4},
                               {2, 5}, {0, 3}, {1, 4}, {2, 4}, {1, 3}, {2, 3}};

template <class S>
static void network_sort(int n, const S* v, int* order) {
  for (int i = 0; i < n; ++i) order[i] = i;
  const int(*net)[2] = nullptr;
  int m = 0;
  if (n == 2) { net = kNet2; m = 1; }
  else if (n == 3) { net = kNet3; m = 3; }
  else if (n == 6) { net = kNet6; m = 12; }
  else if (n > 1) {
    // other sizes (the curvature models' 7 and 30 offsets): insertion sort;
    // the update depends on the order of equal values only through rounding
    for (int i = 1; i < n; ++i)
      for (int j = i; j > 0 && v[order[j]] < v[order[j - 1]]; --j) std::swap(order[j], order[j - 1]);
  }
  for (int c = 0; c < m; ++c) {
    const int i = net[c][0], j = net[c][1];
    const int oi = order[i], oj = order[j];
    const bool b = v[oi] < v[oj];
    order[i] = b ? oi : oj;
    order[j] = b ? oj : oi;
  }
}

// HFM.h:165-207: active-set quadratic.  v[k] are the (already minimised)
// neighbour values, w[k] the weights.  Returns the new value; fills flow
// weights w_k (val - v_k)_+ when fw != nullptr.
template <class S>
static S hfm_solve(int nact, const S* w, const S* v, S rhs, S* fw) {
  int order[kMaxAct];
  network_sort<S>(nact, v, order);
  const S vmin = v[order[0]];
  if (vmin == Inf<S>()) {
    if (fw) for (int k = 0; k < nact; ++k) fw[k] = 0;
    return vmin;
  }
  S a = w[order[0]], b = 0, c = -rhs * rhs;
  for (int kk = 1; kk < nact; ++kk) {
    const int k = order[kk];
    const S t = v[k] - vmin;
    if ((a * t - 2 * b) * t + c > 0) break;
    a += w[k];
    b += w[k] * t;
    c += w[k] * t * t;
  }
  const S delta = b * b - a * c;
  const S sdelta = std::sqrt(delta > S(0) ? delta : S(0));
  const S value = (a == S(0)) ? Inf<S>() : (b + sdelta) / a;
  const S val = vmin + value;
  if (fw) {
    if (val == Inf<S>()) {
      for (int k = 0; k < nact; ++k) fw[k] = 0;
    } else {
      for (int k = 0; k < nact; ++k) {
        const S diff = val - v[k];
        fw[k] = w[k] * (diff > S(0) ? diff : S(0));
      }
    }
  }
  return val;
}

// ---------------------------------------------------------------- grid
struct Grid {
  int d;
  int64_t shape[3];
  int periodic[3];
  int64_t N;
  int64_t stride[3];
  void init(int d_, const int64_t* sh, const int* per) {
    d = d_;
    N = 1;
    for (int k = 0; k < 3; ++k) { shape[k] = 1; periodic[k] = 0; }
    for (int k = 0; k < d; ++k) { shape[k] = sh[k]; periodic[k] = per ? per[k] : 0; N *= sh[k]; }
    int64_t s = 1;
    for (int k = d - 1; k >= 0; --k) { stride[k] = s; s *= shape[k]; }
  }
  void pos(int64_t n, int64_t* x) const {
    for (int k = d - 1; k >= 0; --k) { x[k] = n % shape[k]; n /= shape[k]; }
  }
  // Grid.h:68-80 InRange_per / Index_per: periodic axes wrap, others reject.
  int64_t index(const int64_t* y) const {
    int64_t n = 0;
    for (int k = 0; k < d; ++k) {
      int64_t yk = y[k];
      if (periodic[k]) { yk %= shape[k]; if (yk < 0) yk += shape[k]; }
      else if (yk < 0 || yk >= shape[k]) return -1;
      n += yk * stride[k];
    }
    return n;
  }
};

// Walls.h:13-14: integer division rounding half away from zero (d > 0 here).
static inline int div_round_closest(int n, int d) {
  return ((n < 0) ^ (d < 0)) ? ((n - d / 2) / d) : ((n + d / 2) / d);
}

// Walls.h:18-50: is the segment [x, x+e] free of walls?  wallDist is the
// saturated L1 distance to the walls (255 = far).  The walk samples
// w_k = round(e k / |e|_inf), k = 1..|e|_inf, and concludes as soon as the
// remaining L1 length |e - w_k|_1 is below the wall distance at x + w_k.
static bool visible(const Grid& G, const int64_t* x, const int* e,
                    const uint8_t* wallDist) {
  if (!wallDist) return true;
  const uint8_t wx = wallDist[G.index(x)];
  if (wx == 255) return true;
  int l1 = 0, linf = 0;
  for (int k = 0; k < G.d; ++k) { l1 += std::abs(e[k]); linf = std::max(linf, std::abs(e[k])); }
  if (l1 < wx) return true;
  for (int t = 1; t <= linf; ++t) {
    int64_t y[3];
    int diffl1 = 0;
    for (int k = 0; k < G.d; ++k) {
      const int w = div_round_closest(e[k] * t, linf);
      diffl1 += std::abs(e[k] - w);
      y[k] = x[k] + w;
    }
    const int64_t ny = G.index(y);
    const int value = (ny >= 0) ? wallDist[ny] : 255;
    if (diffl1 < value) return true;
    if (value == 0) return false;
  }
  return true;
}

// Scheme order (HFM.h:91-134): 2 enables the second-order correction where
// the solution is smooth; set for the whole library by orc_set_order.
static int g_order = 1;
static double g_order2_threshold = 0.3;

// Source factorization (EuclideanFactor.h:8-30, Constants.h:50-61,
// FiniteDifferences.h:30-45), set for the whole library by orc_set_factor:
// inside |x - origin|^2 < r2 the value of neighbour x + s e is shifted by
// fact_s = T(x) - T(x + s e) + s <e, grad T(x)>, T(y) = |y - origin|_M, and
// the further neighbour x + 2 s e (order 2) by
// fact2_s = 2 fact_s - (T(x) - 2 T(x + s e) + T(x + 2 s e)).
struct Factor {
  double r2 = 0.;       // 0: off
  double origin[3] = {0., 0., 0.};
  double m[14] = {};  // diagonal, lower triangle, or TTI [l, q, A]
};
static Factor g_factor;

// <x, M y> with one fma per coefficient after the rounded coordinate
// products (off-diagonal pairs fma(x_i, y_j, x_j y_i)).
template <class S>
static S factor_dot(int d, bool diag, const S* m, const S* x, const S* y) {
  S r = 0;
  if (diag) {
    for (int i = 0; i < d; ++i) r = fmaT(x[i] * y[i], m[i], r);
    return r;
  }
  int k = 0;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j <= i; ++j, ++k) {
      const S p = (i == j) ? x[i] * y[i] : fmaT(x[i], y[j], x[j] * y[i]);
      r = fmaT(p, m[k], r);
    }
  return r;
}

// EuclideanFactor.h:8-30: norm differences in cancellation-free form.
template <class S>
static void euclid_factor(S xx, S xe, S ee, S* f, S* f2) {
  const S nx = std::sqrt(xx);
  const S nxm = std::sqrt(xx - S(2) * xe + ee), nxp = std::sqrt(xx + S(2) * xe + ee);
  const S dm = (S(2) * xe - ee) / (nx + nxm);
  const S dp = (-S(2) * xe - ee) / (nx + nxp);
  const S ge = xe / nx;
  f[0] = dm - ge;
  f[1] = ge + dp;
  const S nxm2 = std::sqrt(xx - S(4) * xe + S(4) * ee), nxp2 = std::sqrt(xx + S(4) * xe + S(4) * ee);
  const S am = (-S(2) * xe + S(3) * ee) / (nxm + nxm2);
  const S ap = (S(2) * xe + S(3) * ee) / (nxp + nxp2);
  f2[0] = S(2) * f[0] - (dm + am);
  f2[1] = S(2) * f[1] - (dp + ap);
}

struct Problem {
  int model, d;
  int order;
  double order2_threshold;
  Factor factor;
  Grid G;
  const double* geom;       // ngeom * N, component-major
  const double* rhs;        // N: cost, seed value at seeds, +inf in walls
  const uint8_t* seeds;     // N: fixed cells (seeds and walls)
  const uint8_t* wallDist;  // N or null
  double iso_w[3];
};

// Gather neighbour values for cell n over the sub-stencil starting at
// `base` (HFM.h:39-77, FiniteDifferences.h:20-106).
template <class S, class Get>
static void neighbours(const Problem& P, const Stencil<S>& st, int base, int64_t n, Get get,
                       S* v, int* side, int64_t* nidx, S* v2 = nullptr) {
  int64_t x[3];
  P.G.pos(n, x);
  // factorization: position relative to the origin, inside the ball?
  S xr[3] = {0, 0, 0};
  bool fac = false;
  const bool diag = P.model == ISOTROPIC || P.model == DIAGONAL;
  S fm[14];
  const bool tti = P.model == TTI;
  TTIFactorData<S> td;
  S tnx = 0, tgrad[3] = {0, 0, 0};
  if (P.factor.r2 > 0) {
    S n2 = 0;
    for (int c = 0; c < P.d; ++c) {
      xr[c] = S(x[c]) - S(P.factor.origin[c]);
      n2 = fmaT(xr[c], xr[c], n2);
    }
    fac = n2 < S(P.factor.r2);
    for (int c = 0; c < 14; ++c) fm[c] = S(P.factor.m[c]);
    if (fac && tti) {
      tti_factor_data<S>(fm, td);
      tnx = tti_factor_norm<S>(P.d, fm, td, xr, tgrad);
    }
  }
  for (int k = 0; k < st.nsym; ++k) {
    const int* ek = st.e[base + k];
    S shift = 0;
    if (st.has_drift) {
      for (int c = 0; c < P.d; ++c) shift += st.drift[c] * S(ek[c]);
    }
    S sh[2] = {shift, -shift}, sh2[2] = {0, 0};
    if (fac) {
      const S ef[3] = {S(ek[0]), S(ek[1]), S(ek[2])};
      S f[2], f2[2] = {0, 0};
      if (tti)
        tti_factor_pair<S>(P.d, fm, td, xr, tnx, tgrad, ek, f, P.order == 2 ? f2 : nullptr);
      else
        euclid_factor<S>(factor_dot<S>(P.d, diag, fm, xr, xr), factor_dot<S>(P.d, diag, fm, xr, ef),
                         factor_dot<S>(P.d, diag, fm, ef, ef), f, f2);
      sh[0] = f[0] + sh[0];
      sh[1] = f[1] + sh[1];
      sh2[0] = f2[0];
      sh2[1] = f2[1];
    }
    S best = Inf<S>(), best2 = Inf<S>();
    int bs = 0;
    int64_t bi = -1;
    // a forward offset (the last nfwd of the sub-stencil) has the x + e side only
    const bool fwd = k >= st.nsym - st.nfwd;
    for (int s = fwd ? 1 : 0; s < 2; ++s) {
      const int eps = 2 * s - 1;
      int ee[3] = {eps * ek[0], eps * ek[1], eps * ek[2]};
      int64_t y[3], y2[3];
      for (int c = 0; c < P.d; ++c) { y[c] = x[c] + ee[c]; y2[c] = y[c] + ee[c]; }
      S val = Inf<S>(), val2 = Inf<S>();
      int64_t ny = -1;
      if (visible(P.G, x, ee, P.wallDist)) {
        ny = P.G.index(y);
        if (ny >= 0) val = get(ny) + sh[s];
        // further neighbour x + 2 eps e: factorization shift only
        // (FiniteDifferences.h:75-90, HFM.h:116)
        const int64_t ny2 = P.G.index(y2);
        if (ny2 >= 0) val2 = get(ny2) + sh2[s];
      }
      if (s == 0 || fwd) { best = val; best2 = val2; bs = s; bi = ny; }
      else if (best > val) { best = val; best2 = val2; bs = 1; bi = ny; }
    }
    v[k] = best;
    if (v2) v2[k] = best2;
    if (side) side[k] = bs;
    if (nidx) nidx[k] = bi;
  }
}

// The update of cell n (HFMIter.h:29-56): the upwind update of each
// sub-stencil, keeping the smaller (mix_is_min) or the larger value, with the
// reference's comparison `mix_is_min == (u_new > u_mix)`.  fw/side/nidx (each
// nullable) receive the chosen sub-stencil's flow weights, sides and
// neighbour indices, *kmix its index.
template <class S, class Get>
static S cell_update(const Problem& P, const Stencil<S>& st, int64_t n, Get get, S rhs, S* fw,
                     int* side, int64_t* nidx, int* kmix, S u0 = Inf<S>()) {
  const int ns = st.nsym;
  for (int k = 0; k < ns; ++k) {
    if (fw) fw[k] = 0;
    if (side) side[k] = 0;
    if (nidx) nidx[k] = -1;
  }
  if (kmix) *kmix = 0;
  if (st.nmix == 1 && P.order == 2) {
    // HFM.h:91-134: where |u0 - 2v + v2| < threshold * max_k(u0 - v_k)_+ and
    // u0 > v, use v + (v - v2)/3 with weight 9/4 w (flow weight 3/2 w (u-v)+).
    S v[kMaxAct], v2[kMaxAct], w[kMaxAct];
    bool o2[kMaxAct];
    neighbours<S>(P, st, 0, n, get, v, side, nidx, v2);
    S bound = 0;
    for (int k = 0; k < ns; ++k) bound = std::max(bound, u0 - v[k]);
    const S thr = S(P.order2_threshold);
    for (int k = 0; k < ns; ++k) {
      const S diff2 = std::fabs(u0 - 2 * v[k] + v2[k]);
      o2[k] = diff2 < thr * bound && u0 > v[k];
      w[k] = st.w[k];
      if (o2[k]) {
        v[k] += (v[k] - v2[k]) / 3;
        w[k] *= S(9) / S(4);
      }
    }
    const S val = hfm_solve<S>(ns, w, v, rhs, nullptr);
    if (fw && val != Inf<S>()) {
      for (int k = 0; k < ns; ++k) {
        const S diff = val - v[k];
        fw[k] = st.w[k] * (diff > S(0) ? diff : S(0)) * (o2[k] ? S(1.5) : S(1));
      }
    }
    return val;
  }
  // curvature models: an offset of weight 0 takes no part in the update (a
  // much smaller value of it would otherwise set the base of the quadratic
  // and cancel its FP32 digits); exact arithmetic gives the same value
  const bool drop0 = curvature(P.model);
  if (st.nmix == 1) {
    S v[kMaxAct];
    neighbours<S>(P, st, 0, n, get, v, side, nidx);
    if (drop0)
      for (int k = 0; k < ns; ++k) v[k] = st.w[k] > S(0) ? v[k] : Inf<S>();
    return hfm_solve<S>(ns, st.w, v, rhs, fw);
  }
  S best = st.mix_is_min ? Inf<S>() : -Inf<S>();
  for (int m = 0; m < st.nmix; ++m) {
    S v[kMaxAct], f[kMaxAct], v2[kMaxAct], w[kMaxAct];
    int sd[kMaxAct];
    int64_t ni[kMaxAct];
    neighbours<S>(P, st, m * ns, n, get, v, sd, ni, P.order == 2 ? v2 : nullptr);
    for (int k = 0; k < ns; ++k) w[k] = st.w[m * ns + k];
    if (drop0)
      for (int k = 0; k < ns; ++k) v[k] = w[k] > S(0) ? v[k] : Inf<S>();
    bool o2[kMaxAct];
    for (int k = 0; k < kMaxAct; ++k) o2[k] = false;
    if (P.order == 2) {
      // the second-order correction of each sub-stencil (HFMIter.h:43 passes
      // every sub-stencil's further neighbours to HFMUpdate, HFM.h:91-134)
      S bound = 0;
      for (int k = 0; k < ns; ++k) bound = std::max(bound, u0 - v[k]);
      const S thr = S(P.order2_threshold);
      for (int k = 0; k < ns; ++k) {
        const S diff2 = std::fabs(u0 - 2 * v[k] + v2[k]);
        o2[k] = diff2 < thr * bound && u0 > v[k];
        if (o2[k]) {
          v[k] += (v[k] - v2[k]) / 3;
          w[k] *= S(9) / S(4);
        }
      }
    }
    S val = hfm_solve<S>(ns, w, v, rhs, f);
    if (P.order == 2) {
      for (int k = 0; k < ns; ++k) {
        const S diff = val - v[k];
        f[k] = (val == Inf<S>()) ? S(0)
                                 : st.w[m * ns + k] * (diff > S(0) ? diff : S(0)) * (o2[k] ? S(1.5) : S(1));
      }
    }
    if (st.mix_is_min == (best > val)) {
      best = val;
      for (int k = 0; k < ns; ++k) {
        if (fw) fw[k] = f[k];
        if (side) side[k] = sd[k];
        if (nidx) nidx[k] = ni[k];
      }
      if (kmix) *kmix = m;
    }
  }
  return best;
}

template <class S>
static void stencil_at(const Problem& P, int64_t n, Stencil<S>& st) {
  const int ng = ngeom_of(P.model, P.d);
  S g[16];
  for (int k = 0; k < ng; ++k) g[k] = S(P.geom[k * P.G.N + n]);
  S iw[3] = {S(P.iso_w[0]), S(P.iso_w[1]), S(P.iso_w[2])};
  make_stencil<S>(P.model, P.d, g, iw, st);
}

}  // namespace orc

using namespace orc;

static void fill_problem(Problem& P, int model, int d, const int64_t* shape,
                         const int* periodic, const double* geom, const double* rhs,
                         const uint8_t* seeds, const uint8_t* wallDist,
                         const double* iso_w) {
  P.model = model;
  P.d = d;
  P.order = g_order;
  P.order2_threshold = g_order2_threshold;
  P.factor = g_factor;
  P.G.init(d, shape, periodic);
  P.geom = geom;
  P.rhs = rhs;
  P.seeds = seeds;
  P.wallDist = wallDist;
  for (int k = 0; k < 3; ++k) P.iso_w[k] = (iso_w && k < d) ? iso_w[k] : 0.;
}

// Strict (Jacobi) global iteration with the reference stopping rule
// (Update.h:211-219, DECREASING enabled).  u[N] is in/out (initial guess, a
// supersolution; +inf for a cold start).  Returns the iteration count
// (nitermax if not converged).
template <class S>
static int64_t jacobi(const Problem& P, double atol, double rtol, int64_t nitermax, double* uio) {
  const int64_t N = P.G.N;
  std::vector<Stencil<S>> st(N);
  for (int64_t n = 0; n < N; ++n) stencil_at<S>(P, n, st[n]);
  std::vector<S> u(N), un(N);
  for (int64_t n = 0; n < N; ++n) u[n] = P.seeds[n] ? S(P.rhs[n]) : S(uio[n]);
  int64_t it = 0;
  for (; it < nitermax; ++it) {
    bool changed = false;
    for (int64_t n = 0; n < N; ++n) {
      if (P.seeds[n]) { un[n] = u[n]; continue; }
      const S val = cell_update<S>(P, st[n], n, [&](int64_t m) { return u[m]; }, S(P.rhs[n]),
                                   nullptr, nullptr, nullptr, nullptr, u[n]);
      const S nv = (val < u[n]) ? val : u[n];
      const S diff = std::fabs(u[n] - nv);
      const S tol = S(atol) + S(rtol) * std::fabs(nv);
      if (!(std::isnan(diff) || diff <= tol)) changed = true;
      un[n] = nv;
    }
    u.swap(un);
    if (!changed) break;
  }
  for (int64_t n = 0; n < N; ++n) uio[n] = double(u[n]);
  return it;
}

// Tile adaptive Gauss-Seidel on all host cores (OpenMP): the multi-core CPU
// baseline of the benchmark, the same discrete scheme as the device solver
// (_solvers.py:122-193 on the host).  Tiles of side `tile` along every axis;
// a pass relaxes the active tiles in parallel, each by `niter` in-place
// sweeps in row-major order; a tile whose values changed by more than
// atol + rtol |u| activates itself and its 2d face neighbours for the next
// pass.  Values of other tiles are read and written with relaxed atomic
// accesses: the chaotic relaxation the monotone scheme converges under.
// Stencils are built once per cell.  Returns the passes; *visits = tile
// visits.
template <class S>
static int64_t tiles_omp(const Problem& P, int tile, int niter, double atol, double rtol,
                         int64_t nitermax, double* u, int64_t* visits_out) {
  const int64_t N = P.G.N;
  const int d = P.d;
  std::vector<Stencil<S>> st(N);
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) stencil_at<S>(P, n, st[n]);
  int64_t so[3] = {1, 1, 1}, ntiles = 1;
  for (int c = 0; c < d; ++c) {
    so[c] = (P.G.shape[c] + tile - 1) / tile;
    ntiles *= so[c];
  }
  auto ld = [&](int64_t m) {
    double r;
    __atomic_load(&u[m], &r, __ATOMIC_RELAXED);
    return S(r);
  };
  for (int64_t n = 0; n < N; ++n)
    if (P.seeds[n]) u[n] = P.rhs[n];
  std::vector<uint8_t> next(ntiles, 0);
  std::vector<int64_t> list;
  // initial list: tiles holding a seed
  for (int64_t n = 0; n < N; ++n)
    if (P.seeds[n] && P.rhs[n] < Inf<double>()) {
      int64_t x[3];
      P.G.pos(n, x);
      int64_t t = 0;
      for (int c = 0; c < d; ++c) t = t * so[c] + x[c] / tile;
      next[t] = 1;
    }
  int64_t pass = 0, visits = 0;
  for (; pass < nitermax; ++pass) {
    list.clear();
    for (int64_t t = 0; t < ntiles; ++t)
      if (next[t]) {
        list.push_back(t);
        next[t] = 0;
      }
    if (list.empty()) break;
    visits += int64_t(list.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t li = 0; li < int64_t(list.size()); ++li) {
      const int64_t t = list[li];
      int64_t to[3] = {0, 0, 0}, rem = t;
      for (int c = d - 1; c >= 0; --c) {
        to[c] = rem % so[c];
        rem /= so[c];
      }
      int64_t lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
      for (int c = 0; c < d; ++c) {
        lo[c] = to[c] * tile;
        hi[c] = std::min<int64_t>(lo[c] + tile, P.G.shape[c]);
      }
      bool any = false;
      for (int it = 0; it < niter; ++it) {
        bool swept = false;
        int64_t x[3];
        for (x[0] = lo[0]; x[0] < hi[0]; ++x[0])
          for (x[1] = lo[1]; x[1] < hi[1]; ++x[1])
            for (x[2] = lo[2]; x[2] < hi[2]; ++x[2]) {
              const int64_t n = P.G.index(x);
              if (P.seeds[n]) continue;
              const S un = ld(n);
              const S val = cell_update<S>(P, st[n], n, ld, S(P.rhs[n]), nullptr, nullptr,
                                           nullptr, nullptr, un);
              if (val < un) {
                const double vd = double(val);
                __atomic_store(&u[n], &vd, __ATOMIC_RELAXED);
                const S tol = S(atol) + S(rtol) * std::fabs(val);
                if (!(un - val <= tol)) swept = any = true;
              }
            }
        if (!swept) break;
      }
      if (any) {
        const uint8_t one = 1;
        __atomic_store(&next[t], &one, __ATOMIC_RELAXED);
        for (int c = 0; c < d; ++c)
          for (int s2 = -1; s2 <= 1; s2 += 2) {
            int64_t y[3] = {to[0], to[1], to[2]};
            y[c] += s2;
            if (P.G.periodic[c]) y[c] = (y[c] + so[c]) % so[c];  // Propagation.h:25-28
            else if (y[c] < 0 || y[c] >= so[c]) continue;
            int64_t tn = 0;
            for (int k = 0; k < d; ++k) tn = tn * so[k] + y[k];
            __atomic_store(&next[tn], &one, __ATOMIC_RELAXED);
          }
      }
    }
  }
  *visits_out = visits;
  return pass;
}

// The TTI factorization at n points x[n][d] (relative to the origin) for the
// offsets e[n][3]: norm[n], grad[n][d], fact[n][2], fact2[n][2] (each output
// nullable), in FP32 (f32 = 1) or FP64.
template <class S>
static void tti_factor_points(int d, const double* fm64, int64_t n, const double* x, const int32_t* e,
                              double* norm, double* grad, double* fact, double* fact2) {
  S fm[14];
  for (int k = 0; k < 14; ++k) fm[k] = S(fm64[k]);
  TTIFactorData<S> t;
  tti_factor_data<S>(fm, t);
  for (int64_t i = 0; i < n; ++i) {
    S xs[3] = {0, 0, 0}, g[3] = {0, 0, 0};
    for (int c = 0; c < d; ++c) xs[c] = S(x[i * d + c]);
    const S nx = tti_factor_norm<S>(d, fm, t, xs, g);
    if (norm) norm[i] = double(nx);
    if (grad)
      for (int c = 0; c < d; ++c) grad[i * d + c] = double(g[c]);
    if (fact && e) {
      S f[2], f2[2];
      tti_factor_pair<S>(d, fm, t, xs, nx, g, e + 3 * i, f, fact2 ? f2 : nullptr);
      fact[2 * i] = double(f[0]);
      fact[2 * i + 1] = double(f[1]);
      if (fact2) {
        fact2[2 * i] = double(f2[0]);
        fact2[2 * i + 1] = double(f2[1]);
      }
    }
  }
}

extern "C" {

int orc_version() { return 8; }

// Relaxation eps^2 of the curvature models' self outer products.
void orc_set_curvature_relax(double relax) { g_curvature_relax = relax; }

// 1: the Convex variants of ReedsSheppForward2, Elastica2 and Dubins2.
void orc_set_curvature_convex(int convex) { g_curvature_convex = convex != 0; }

// Node l of the n-point Fejer rule: out = (cos phi_l, sin phi_l, weight).
void orc_fejer(int l, int n, double* out) { fejer(l, n, out[0], out[1], out[2]); }

// Threads the OpenMP solver runs on (omp_get_max_threads).
int orc_omp_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int64_t orc_solve_tiles_omp(int model, int d, const int64_t* shape, const int* periodic,
                            const double* geom, const double* rhs, const uint8_t* seeds,
                            const uint8_t* wallDist, const double* iso_w, int tile, int niter,
                            double atol, double rtol, int64_t nitermax, int64_t* visits,
                            double* u) {
  if (g_order != 1 || g_factor.r2 > 0) return -1;
  Problem P;
  fill_problem(P, model, d, shape, periodic, geom, rhs, seeds, wallDist, iso_w);
  for (int64_t n = 0; n < P.G.N; ++n) u[n] = Inf<double>();
  return tiles_omp<double>(P, tile, niter, atol, rtol, nitermax, u, visits);
}

void orc_set_order(int order, double threshold) {
  g_order = order;
  g_order2_threshold = threshold;
}

// Source factorization for the following solves (r2 = 0 switches it off):
// origin[3] in grid units, m[14] the metric at the origin in grid units.
void orc_set_factor(double r2, const double* origin, const double* m) {
  g_factor = Factor{};
  g_factor.r2 = r2;
  if (r2 > 0) {
    for (int k = 0; k < 3; ++k) g_factor.origin[k] = origin[k];
    for (int k = 0; k < 14; ++k) g_factor.m[k] = m[k];
  }
}

void orc_tti_factor(int d, int f32, const double* fm, int64_t n, const double* x, const int32_t* e,
                    double* norm, double* grad, double* fact, double* fact2) {
  if (f32) tti_factor_points<float>(d, fm, n, x, e, norm, grad, fact, fact2);
  else tti_factor_points<double>(d, fm, n, x, e, norm, grad, fact, fact2);
}

// Selling decomposition of n symmetric matrices (row n: symdim coefficients,
// lower triangle (d00, d10, d11, d20, d21, d22)).  Outputs weights[n][symdim]
// and offsets[n][symdim][d].  use_float selects FP32 arithmetic.
void orc_selling(int d, int64_t n, const double* m, double* w, int32_t* e, int use_float) {
  const int sd = symdim(d);
  for (int64_t c = 0; c < n; ++c) {
    int ee[6][3];
    if (use_float) {
      float mf[6], wf[6];
      for (int k = 0; k < sd; ++k) mf[k] = float(m[c * sd + k]);
      if (d == 2) selling2<float>(mf, wf, ee); else selling3<float>(mf, wf, ee);
      for (int k = 0; k < sd; ++k) w[c * sd + k] = wf[k];
    } else {
      if (d == 2) selling2<double>(m + c * sd, w + c * sd, ee);
      else selling3<double>(m + c * sd, w + c * sd, ee);
    }
    for (int k = 0; k < sd; ++k)
      for (int l = 0; l < d; ++l) e[(c * sd + k) * d + l] = ee[k][l];
  }
}

// The upwind update for n independent cells: w[n][nact], v[n][nact] (already
// minimised over +-e), rhs[n] -> u[n], flow weights fw[n][nact] (nullable).
void orc_hfm_update(int nact, int64_t n, const double* w, const double* v,
                    const double* rhs, double* u, double* fw, int use_float) {
  if (nact < 1 || nact > kMaxAct) return;
  for (int64_t c = 0; c < n; ++c) {
    if (use_float) {
      float wf[kMaxAct], vf[kMaxAct], f[kMaxAct];
      for (int k = 0; k < nact; ++k) { wf[k] = float(w[c * nact + k]); vf[k] = float(v[c * nact + k]); }
      u[c] = hfm_solve<float>(nact, wf, vf, float(rhs[c]), fw ? f : nullptr);
      if (fw) for (int k = 0; k < nact; ++k) fw[c * nact + k] = f[k];
    } else {
      u[c] = hfm_solve<double>(nact, w + c * nact, v + c * nact, rhs[c],
                               fw ? fw + c * nact : nullptr);
    }
  }
}

// Per-cell stencil export for a whole problem, FP32 or FP64: weights
// w[N][nmix*nsym], offsets e[N][nmix*nsym][d], mix flag mix[N] (nullable;
// 1 = min of the sub-stencil updates).
void orc_stencils(int model, int d, const int64_t* shape, const double* geom,
                  const double* iso_w, int use_float, double* w, int32_t* e, int32_t* mix) {
  Problem P;
  fill_problem(P, model, d, shape, nullptr, geom, nullptr, nullptr, nullptr, iso_w);
  const int ns = nsym_of(model, d) * nmix_of(model);
  for (int64_t n = 0; n < P.G.N; ++n) {
    static thread_local int ee[kMaxSlots][3];
    bool mn = true;
    if (use_float) {
      Stencil<float> st;
      stencil_at<float>(P, n, st);
      mn = st.mix_is_min;
      for (int k = 0; k < ns; ++k) { w[n * ns + k] = st.w[k]; for (int l = 0; l < d; ++l) ee[k][l] = st.e[k][l]; }
    } else {
      Stencil<double> st;
      stencil_at<double>(P, n, st);
      mn = st.mix_is_min;
      for (int k = 0; k < ns; ++k) { w[n * ns + k] = st.w[k]; for (int l = 0; l < d; ++l) ee[k][l] = st.e[k][l]; }
    }
    for (int k = 0; k < ns; ++k)
      for (int l = 0; l < d; ++l) e[(n * ns + k) * d + l] = ee[k][l];
    if (mix) mix[n] = mn ? 1 : 0;
  }
}

// TTI_.h:76-107 for n cells: l[n][2], q[n][3] -> diag[n][nmix][2], returns
// mix_is_min per cell in mix[n].  FP32 or FP64.
void orc_tti_diags(int64_t n, const double* l, const double* q, double* diag, int32_t* mix,
                   int use_float) {
  for (int64_t c = 0; c < n; ++c) {
    if (use_float) {
      float lf[2] = {float(l[2 * c]), float(l[2 * c + 1])};
      float qf[3] = {float(q[3 * c]), float(q[3 * c + 1]), float(q[3 * c + 2])};
      float ds[kNMix][2];
      mix[c] = tti_diags<float>(lf, qf, ds);
      for (int i = 0; i < kNMix; ++i) { diag[(c * kNMix + i) * 2] = ds[i][0]; diag[(c * kNMix + i) * 2 + 1] = ds[i][1]; }
    } else {
      double ds[kNMix][2];
      mix[c] = tti_diags<double>(l + 2 * c, q + 3 * c, ds);
      for (int i = 0; i < kNMix; ++i) { diag[(c * kNMix + i) * 2] = ds[i][0]; diag[(c * kNMix + i) * 2 + 1] = ds[i][1]; }
    }
  }
}

// Exact fixed point by fast marching (FP64).  u[N] receives the values.
// Returns the number of accepted cells.  With u0 (nullable), the fixed point
// of u = min(u0, scheme(u)) (seeds keep their value): the non-seed cells start
// as tentative values u0, which is where the decreasing iteration started
// from u0 ends (the continuation of a solve after a paste between charts).
static int64_t solve_fm(int model, int d, const int64_t* shape, const int* periodic, const double* geom,
                        const double* rhs, const uint8_t* seeds, const uint8_t* wallDist,
                        const double* iso_w, const double* u0, double* u) {
  // the second-order and factored schemes are solved by iteration
  if (g_order != 1 || g_factor.r2 > 0) return -1;
  Problem P;
  fill_problem(P, model, d, shape, periodic, geom, rhs, seeds, wallDist, iso_w);
  const int64_t N = P.G.N;
  const int ns = nsym_of(model, d) * nmix_of(model);  // offsets of all sub-stencils
  // Stencils and reverse adjacency (cells whose update reads a given cell).
  std::vector<Stencil<double>> st(N);
  std::vector<int64_t> deg(N + 1, 0);
  std::vector<int64_t> nbr((size_t)N * 2 * ns, -1);
  for (int64_t n = 0; n < N; ++n) {
    stencil_at<double>(P, n, st[n]);
    int64_t x[3];
    P.G.pos(n, x);
    for (int k = 0; k < ns; ++k)
      for (int s = 0; s < 2; ++s) {
        const int eps = 2 * s - 1;
        int ee[3] = {eps * st[n].e[k][0], eps * st[n].e[k][1], eps * st[n].e[k][2]};
        int64_t y[3];
        for (int c = 0; c < d; ++c) y[c] = x[c] + ee[c];
        const int64_t ny = visible(P.G, x, ee, wallDist) ? P.G.index(y) : -1;
        nbr[(size_t)n * 2 * ns + 2 * k + s] = ny;
        if (ny >= 0) deg[ny + 1]++;
      }
  }
  for (int64_t n = 0; n < N; ++n) deg[n + 1] += deg[n];
  std::vector<int64_t> rev(deg[N]);
  {
    std::vector<int64_t> cur(deg.begin(), deg.end() - 1);
    for (int64_t n = 0; n < N; ++n)
      for (int j = 0; j < 2 * ns; ++j) {
        const int64_t ny = nbr[(size_t)n * 2 * ns + j];
        if (ny >= 0) rev[cur[ny]++] = n;
      }
  }
  std::vector<uint8_t> acc(N, 0);
  for (int64_t n = 0; n < N; ++n) u[n] = Inf<double>();
  typedef std::pair<double, int64_t> QE;
  std::priority_queue<QE, std::vector<QE>, std::greater<QE>> heap;
  for (int64_t n = 0; n < N; ++n)
    if (seeds[n]) {
      u[n] = rhs[n];
      if (u[n] < Inf<double>()) heap.push(QE(u[n], n));
      else acc[n] = 1;  // walls: fixed at +inf
    } else if (u0 && u0[n] < Inf<double>()) {
      u[n] = u0[n];
      heap.push(QE(u[n], n));
    }
  int64_t naccepted = 0;
  auto get_acc = [&](int64_t m) { return acc[m] ? u[m] : Inf<double>(); };
  while (!heap.empty()) {
    const QE top = heap.top();
    heap.pop();
    const int64_t n = top.second;
    if (acc[n] || top.first > u[n]) continue;
    acc[n] = 1;
    ++naccepted;
    for (int64_t r = deg[n]; r < deg[n + 1]; ++r) {
      const int64_t m = rev[r];
      if (acc[m] || seeds[m]) continue;
      const double val = cell_update<double>(P, st[m], m, get_acc, rhs[m], nullptr, nullptr,
                                             nullptr, nullptr);
      if (val < u[m]) {
        u[m] = val;
        heap.push(QE(val, m));
      }
    }
  }
  return naccepted;
}

int64_t orc_solve_fm(int model, int d, const int64_t* shape, const int* periodic,
                     const double* geom, const double* rhs, const uint8_t* seeds,
                     const uint8_t* wallDist, const double* iso_w, double* u) {
  return solve_fm(model, d, shape, periodic, geom, rhs, seeds, wallDist, iso_w, nullptr, u);
}

int64_t orc_solve_fm_from(int model, int d, const int64_t* shape, const int* periodic,
                          const double* geom, const double* rhs, const uint8_t* seeds,
                          const uint8_t* wallDist, const double* iso_w, const double* u0, double* u) {
  return solve_fm(model, d, shape, periodic, geom, rhs, seeds, wallDist, iso_w, u0, u);
}

int64_t orc_solve_jacobi(int model, int d, const int64_t* shape, const int* periodic,
                         const double* geom, const double* rhs, const uint8_t* seeds,
                         const uint8_t* wallDist, const double* iso_w, double atol,
                         double rtol, int64_t nitermax, int use_float, double* u) {
  Problem P;
  fill_problem(P, model, d, shape, periodic, geom, rhs, seeds, wallDist, iso_w);
  return use_float ? jacobi<float>(P, atol, rtol, nitermax, u)
                   : jacobi<double>(P, atol, rtol, nitermax, u);
}

// Flow export from converged values (Update.h:171-209 with FLOW_* on): one
// update per non-seed cell; writes the updated value back (the reference's
// flow pass is one more global sweep that also stores u), the flow weights
// fw[N][nsym], neighbour linear indices nidx[N][nsym] (-1 when the weight is
// zero), flow vector fv[d][N] = sum_k fw_k (+-e_k), and the weight sum.
int orc_flow(int model, int d, const int64_t* shape, const int* periodic, const double* geom,
             const double* rhs, const uint8_t* seeds, const uint8_t* wallDist,
             const double* iso_w, int use_float, double* u, double* fw, int64_t* nidx,
             double* fv, double* wsum) {
  Problem P;
  fill_problem(P, model, d, shape, periodic, geom, rhs, seeds, wallDist, iso_w);
  const int64_t N = P.G.N;
  const int ns = nsym_of(model, d);
  std::vector<double> u0(u, u + N);
  for (int64_t n = 0; n < N; ++n) {
    double f[kMaxAct];
    int side[kMaxAct];
    int64_t ni[kMaxAct];
    int e[kMaxAct][3];
    for (int k = 0; k < kMaxAct; ++k) {
      f[k] = 0;
      side[k] = 0;
      ni[k] = -1;
    }
    if (!seeds[n]) {
      int km = 0;
      if (use_float) {
        Stencil<float> st;
        stencil_at<float>(P, n, st);
        float ff[kMaxAct];
        const float val = cell_update<float>(P, st, n, [&](int64_t m) { return float(u0[m]); },
                                             float(rhs[n]), ff, side, ni, &km, float(u0[n]));
        if (val < float(u0[n])) u[n] = val;
        for (int k = 0; k < ns; ++k) {
          f[k] = ff[k];
          for (int c = 0; c < 3; ++c) e[k][c] = st.e[km * ns + k][c];
        }
      } else {
        Stencil<double> st;
        stencil_at<double>(P, n, st);
        const double val = cell_update<double>(P, st, n, [&](int64_t m) { return u0[m]; }, rhs[n],
                                               f, side, ni, &km, u0[n]);
        if (val < u0[n]) u[n] = val;
        for (int k = 0; k < ns; ++k) for (int c = 0; c < 3; ++c) e[k][c] = st.e[km * ns + k][c];
      }
    } else {
      for (int k = 0; k < ns; ++k) for (int c = 0; c < 3; ++c) e[k][c] = 0;
    }
    double s = 0;
    double vec[3] = {0, 0, 0};
    for (int k = 0; k < ns; ++k) {
      const int eps = 2 * side[k] - 1;
      fw[n * ns + k] = f[k];
      nidx[n * ns + k] = (f[k] != 0) ? ni[k] : -1;
      s += f[k];
      for (int c = 0; c < d; ++c) vec[c] += f[k] * double(eps * e[k][c]);
    }
    wsum[n] = s;
    for (int c = 0; c < d; ++c) fv[c * N + n] = vec[c];
  }
  return 0;
}

// Reverse-mode sensitivity: solves (D - W)^T lam = sens exactly (FP64) by
// visiting cells in decreasing order of u (the upwind graph is acyclic in that
// order: a positive flow weight from x to y implies u(x) > u(y)).  D = sum of
// flow weights (1 where it vanishes: seeds, walls).  nrhs right-hand sides,
// sens/lam layout [nrhs][N].
int orc_adjoint(int64_t N, int nact, const double* u, const double* fw, const int64_t* nidx,
                int nrhs, const double* sens, double* lam) {
  std::vector<int64_t> order;
  order.reserve(N);
  for (int64_t n = 0; n < N; ++n)
    if (std::isfinite(u[n])) order.push_back(n);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return u[a] > u[b]; });
  for (int r = 0; r < nrhs; ++r) {
    double* L = lam + (int64_t)r * N;
    const double* Sv = sens + (int64_t)r * N;
    std::vector<double> acc(N, 0.);
    for (int64_t n = 0; n < N; ++n) L[n] = 0.;
    for (int64_t n : order) {
      double diag = 0;
      for (int k = 0; k < nact; ++k) diag += fw[n * nact + k];
      if (diag == 0) diag = 1;
      const double l = (Sv[n] + acc[n]) / diag;
      L[n] = l;
      for (int k = 0; k < nact; ++k) {
        const double w = fw[n * nact + k];
        const int64_t m = nidx[n * nact + k];
        if (w != 0 && m >= 0) acc[m] += w * l;
      }
    }
  }
  return 0;
}

// Forward-mode variation: solves (D - W) du = r exactly (FP64) by visiting
// cells in increasing order of u (upwind neighbours come first), the
// untransposed counterpart of orc_adjoint (_PostProcess.py:181-191).
// r/du layout [nrhs][N]; cells with a non-finite u get du = 0.
int orc_forward_ad(int64_t N, int nact, const double* u, const double* fw, const int64_t* nidx,
                   int nrhs, const double* r, double* du) {
  std::vector<int64_t> order;
  order.reserve(N);
  for (int64_t n = 0; n < N; ++n)
    if (std::isfinite(u[n])) order.push_back(n);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return u[a] < u[b]; });
  for (int q = 0; q < nrhs; ++q) {
    double* X = du + (int64_t)q * N;
    const double* R = r + (int64_t)q * N;
    for (int64_t n = 0; n < N; ++n) X[n] = 0.;
    for (int64_t n : order) {
      double diag = 0, acc = R[n];
      for (int k = 0; k < nact; ++k) {
        const double w = fw[n * nact + k];
        const int64_t m = nidx[n * nact + k];
        diag += w;
        if (w != 0 && m >= 0) acc += w * X[m];
      }
      X[n] = acc / (diag == 0 ? 1. : diag);
    }
  }
  return 0;
}

// L1 distance to walls, saturated (_SetGeometry.py:262-282): 0 on walls,
// k for cells at L1 distance k <= bound, 255 beyond the bound.
void orc_wall_distance(int d, const int64_t* shape, const int* periodic, const uint8_t* walls,
                       int bound, uint8_t* dist) {
  Grid G;
  G.init(d, shape, periodic);
  std::vector<int> cur(G.N), nxt(G.N);
  for (int64_t n = 0; n < G.N; ++n) cur[n] = walls[n] ? 0 : 255;
  for (int it = 0; it < bound; ++it) {
    for (int64_t n = 0; n < G.N; ++n) {
      int best = cur[n];
      int64_t x[3];
      G.pos(n, x);
      for (int k = 0; k < d; ++k)
        for (int s = -1; s <= 1; s += 2) {
          int64_t y[3] = {x[0], x[1], x[2]};
          y[k] += s;
          const int64_t m = G.index(y);
          if (m >= 0) best = std::min(best, std::min(255, cur[m] + 1));
        }
      nxt[n] = best;
    }
    cur.swap(nxt);
  }
  for (int64_t n = 0; n < G.N; ++n) dist[n] = (uint8_t)(cur[n] > bound ? 255 : cur[n]);
}

// Visibility of the segments [x_i, x_i + e_i] (n queries, cell-major grid
// positions and offsets) under the walk of Walls.h:18-50.
void orc_visible(int d, const int64_t* shape, const int* periodic, const uint8_t* wall_dist,
                 int64_t n, const int64_t* x, const int32_t* e, uint8_t* out) {
  Grid G;
  G.init(d, shape, periodic);
  for (int64_t q = 0; q < n; ++q) {
    int64_t xx[3] = {0, 0, 0};
    int ee[3] = {0, 0, 0};
    for (int k = 0; k < d; ++k) { xx[k] = x[q * d + k]; ee[k] = e[q * d + k]; }
    out[q] = visible(G, xx, ee, wall_dist) ? 1 : 0;
  }
}

}  // extern "C"
