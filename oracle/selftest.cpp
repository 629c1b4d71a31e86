// selftest.cpp -- drives the oracle's C entry points on small problems of
// every model, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
// (`make -C oracle selftest`).  Test infrastructure only: it checks that the
// restatement reads and writes inside its buffers and is free of undefined
// behaviour, plus a few cheap consistencies (fast marching, the iteration and
// the tile solver reach the same values; the flow weights are non-negative;
// the adjoint of a zero right-hand side is zero).  Exit status 0 on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

extern "C" {
int orc_version();
void orc_set_order(int order, double threshold);
void orc_set_factor(double r2, const double* origin, const double* m);
int64_t orc_solve_fm(int model, int d, const int64_t* shape, const int* periodic, const double* geom,
                     const double* rhs, const uint8_t* seeds, const uint8_t* wallDist,
                     const double* iso_w, double* u);
int64_t orc_solve_jacobi(int model, int d, const int64_t* shape, const int* periodic,
                         const double* geom, const double* rhs, const uint8_t* seeds,
                         const uint8_t* wallDist, const double* iso_w, double atol, double rtol,
                         int64_t nitermax, int use_float, double* u);
int64_t orc_solve_tiles_omp(int model, int d, const int64_t* shape, const int* periodic,
                            const double* geom, const double* rhs, const uint8_t* seeds,
                            const uint8_t* wallDist, const double* iso_w, int tile, int niter,
                            double atol, double rtol, int64_t nitermax, int64_t* visits, double* u);
int orc_flow(int model, int d, const int64_t* shape, const int* periodic, const double* geom,
             const double* rhs, const uint8_t* seeds, const uint8_t* wallDist, const double* iso_w,
             int use_float, double* u, double* fw, int64_t* nidx, double* fv, double* wsum);
int orc_adjoint(int64_t N, int nact, const double* u, const double* fw, const int64_t* nidx, int nrhs,
                const double* sens, double* lam);
void orc_wall_distance(int d, const int64_t* shape, const int* periodic, const uint8_t* walls,
                       int bound, uint8_t* out);
void orc_tti_factor(int d, int f32, const double* fm, int64_t n, const double* x, const int32_t* e,
                    double* norm, double* grad, double* fact, double* fact2);
}

namespace {

enum { ISO = 0, DIAG = 1, RIEM = 2, RAND = 3, TTI = 4, RS2 = 5, RSF2 = 6, ELA2 = 7, DUB2 = 8 };
int g_fail = 0;

void check(bool ok, const char* what, double v = 0.) {
  if (!ok) {
    std::fprintf(stderr, "selftest: FAILED %s (%g)\n", what, v);
    ++g_fail;
  }
}

double max_rel_diff(const std::vector<double>& a, const std::vector<double>& b) {
  double m = 0., s = 0.;
  for (size_t i = 0; i < a.size(); ++i) {
    if (std::isfinite(b[i])) {
      m = std::fmax(m, std::fabs(a[i] - b[i]));
      s = std::fmax(s, std::fabs(b[i]));
    } else if (std::isfinite(a[i])) {
      return INFINITY;
    }
  }
  return s > 0. ? m / s : m;
}

// Geometry of a smooth anisotropic field of the model (component-major).
std::vector<double> geometry(int model, int d, int64_t N, unsigned seed) {
  std::srand(seed);
  auto r = [] { return double(std::rand()) / RAND_MAX; };
  const int sd = d * (d + 1) / 2;
  std::vector<double> g;
  if (model == DIAG) {
    g.resize(size_t(d) * N);
    for (auto& v : g) v = 0.5 + r();
  } else if (model == RIEM || model == RAND) {
    const int ng = sd + (model == RAND ? d : 0);
    g.assign(size_t(ng) * N, 0.);
    for (int64_t n = 0; n < N; ++n) {
      // dual metric: diagonal dominant SPD
      for (int i = 0, k = 0; i < d; ++i)
        for (int j = 0; j <= i; ++j, ++k) g[size_t(k) * N + n] = (i == j) ? 1. + r() : 0.3 * (r() - 0.5);
      if (model == RAND)
        for (int c = 0; c < d; ++c) g[size_t(sd + c) * N + n] = 0.2 * (r() - 0.5);
    }
  } else if (model >= RS2) {
    // curvature models on (x, y, theta): [1/xi, kappa, cos theta, sin theta]
    g.assign(size_t(4) * N, 0.);
    const int64_t nt = 7;  // angles along the last axis (shape[2])
    for (int64_t n = 0; n < N; ++n) {
      const double th = 2. * 3.141592653589793 * double(n % nt) / double(nt);
      g[size_t(n)] = 1. + r();
      g[size_t(N + n)] = 0.3 * (r() - 0.5);
      g[size_t(2 * N + n)] = std::cos(th);
      g[size_t(3 * N + n)] = std::sin(th);
    }
  } else if (model == TTI) {
    const int ng = 5 + d * d;
    g.assign(size_t(ng) * N, 0.);
    for (int64_t n = 0; n < N; ++n) {
      // hexagonal medium (15, 4, 10, 3): l = (18, 13), q = (-90, -110, -60)
      const double l0 = 18., l1 = 13., q00 = -90., q10 = 16. - 150. + 24., q11 = -60.;
      const double vals[5] = {l0, l1, q00, q10, q11};
      for (int k = 0; k < 5; ++k) g[size_t(k) * N + n] = vals[k];
      for (int i = 0; i < d; ++i) g[size_t(5 + i * d + i) * N + n] = 1. + 0.1 * r();
    }
  }
  return g;
}

void run_model(int model, int d, bool walls) {
  const int64_t shape[3] = {d == 2 ? 23 : 9, d == 2 ? 17 : 8, d == 2 ? 1 : 7};
  const int periodic[3] = {0, (d == 3 && model < RS2) ? 1 : 0, model >= RS2 ? 1 : 0};
  int64_t N = 1;
  for (int k = 0; k < d; ++k) N *= shape[k];
  std::vector<double> geom = geometry(model, d, N, 17u + unsigned(model) * 3u + unsigned(d));
  std::vector<double> rhs(size_t(N), 1.);
  std::vector<uint8_t> seeds(size_t(N), 0), wall(size_t(N), 0), wd;
  seeds[size_t(N / 3)] = 1;
  rhs[size_t(N / 3)] = 0.;
  if (walls) {
    for (int64_t n = 0; n < N; ++n)
      if (n % 7 == 3 && n != N / 3) {
        wall[size_t(n)] = 1;
        seeds[size_t(n)] = 1;
        rhs[size_t(n)] = INFINITY;
      }
    wd.resize(size_t(N));
    orc_wall_distance(d, shape, periodic, wall.data(), 6, wd.data());
  }
  const uint8_t* wdp = walls ? wd.data() : nullptr;
  const double iso_w[3] = {1., 1.2, 0.9};
  const double* gp = geom.empty() ? nullptr : geom.data();
  const size_t NN = static_cast<size_t>(N);
  std::vector<double> u_fm(NN), u_j(NN), u_t(NN);
  char what[128];
  std::snprintf(what, sizeof what, "model %d, %dD%s", model, d, walls ? ", walls" : "");
  std::string w0 = std::string(what) + ": fast marching";
  check(orc_solve_fm(model, d, shape, periodic, gp, rhs.data(), seeds.data(), wdp, iso_w, u_fm.data()) >= 0,
        w0.c_str());
  std::fill(u_j.begin(), u_j.end(), INFINITY);  // the decreasing iteration starts from +inf
  orc_solve_jacobi(model, d, shape, periodic, gp, rhs.data(), seeds.data(), wdp, iso_w, 1e-12, 0., 100000, 0,
                   u_j.data());
  const std::string w1 = std::string(what) + ": iteration against fast marching";
  // (fast marching is exact for the symmetric schemes; with a drift it
  // accepts cells slightly out of causal order)
  check(max_rel_diff(u_j, u_fm) < (model == RAND ? 1e-3 : 1e-9), w1.c_str(), max_rel_diff(u_j, u_fm));
  int64_t visits = 0;
  if (orc_solve_tiles_omp(model, d, shape, periodic, gp, rhs.data(), seeds.data(), wdp, iso_w, 4, 4, 1e-12,
                          0., 100000, &visits, u_t.data()) >= 0)
    check(max_rel_diff(u_t, u_j) < 1e-9, (std::string(what) + ": tiles against iteration").c_str(),
          max_rel_diff(u_t, u_j));
  const int nsym = (model == ISO || model == DIAG) ? d
                   : model == RS2 ? 6 : model == RSF2 ? 7 : model == ELA2 ? 30 : model == DUB2 ? 6
                                                                                : d * (d + 1) / 2;
  std::vector<double> fw(NN * nsym), fv(size_t(d) * NN), ws(NN);
  std::vector<int64_t> nidx(NN * nsym);
  std::vector<double> u = u_fm;
  orc_flow(model, d, shape, periodic, gp, rhs.data(), seeds.data(), wdp, iso_w, 0, u.data(), fw.data(),
           nidx.data(), fv.data(), ws.data());
  bool nonneg = true;
  for (double w : fw) nonneg = nonneg && !(w < 0.);
  check(nonneg, (std::string(what) + ": flow weights").c_str());
  std::vector<double> sens(size_t(N), 0.), lam(size_t(N), 1.);
  orc_adjoint(N, nsym, u.data(), fw.data(), nidx.data(), 1, sens.data(), lam.data());
  bool zero = true;
  for (double l : lam) zero = zero && l == 0.;
  check(zero, (std::string(what) + ": zero adjoint").c_str());
  if (model >= RS2) return;  // first order, no factorization
  // second order and factorization, by iteration
  orc_set_order(2, 0.3);
  std::fill(u_j.begin(), u_j.end(), INFINITY);
  orc_solve_jacobi(model, d, shape, periodic, gp, rhs.data(), seeds.data(), wdp, iso_w, 1e-10, 0., 100000, 1,
                   u_j.data());
  orc_set_order(1, 0.3);
  const double origin[3] = {double((N / 3) / (d == 2 ? shape[1] : shape[1] * shape[2])), 1., 1.};
  double fm[14] = {1., 0., 1., 0., 0., 1., 0., 0., 0., 0., 0., 0., 0., 0.};
  if (model == TTI) {
    const double t[14] = {18., 13., -90., -110., -60., 1., 0., 0., 0., 1., 0., 0., 0., 1.};
    for (int k = 0; k < 14; ++k) fm[k] = t[k];
    if (d == 2) { fm[5] = 1.; fm[6] = 0.; fm[7] = 0.; fm[8] = 1.; }
  }
  orc_set_factor(16., origin, fm);
  std::fill(u_j.begin(), u_j.end(), INFINITY);
  orc_solve_jacobi(model, d, shape, periodic, gp, rhs.data(), seeds.data(), wdp, iso_w, 1e-10, 0., 100000, 0,
                   u_j.data());
  orc_set_factor(0., nullptr, nullptr);
}

}  // namespace

int main() {
  check(orc_version() >= 8, "oracle version");
  for (int d = 2; d <= 3; ++d)
    for (int model = ISO; model <= TTI; ++model) {
      run_model(model, d, false);
      run_model(model, d, true);
    }
  for (int model = RS2; model <= DUB2; ++model) {
    run_model(model, 3, false);
    run_model(model, 3, true);
  }
  // the TTI factor norm on a few points, both precisions
  const double fm[14] = {18., 13., -90., -110., -60., 0.9, 0.2, 0., -0.2, 0.9, 0., 0., 0., 1.};
  const double x[6] = {3., -1., 2., 0.5, -4., 2.};
  const int32_t e[6] = {1, 0, 0, 1, -1, 1};
  double nrm[2], grad[6], fact[4], fact2[4];
  for (int f32 = 0; f32 < 2; ++f32) {
    orc_tti_factor(3, f32, fm, 2, x, e, nrm, grad, fact, fact2);
    check(std::isfinite(nrm[0]) && std::isfinite(nrm[1]) && nrm[0] > 0., "tti factor norm");
  }
  if (g_fail) return 1;
  std::printf("selftest: ok\n");
  return 0;
}
