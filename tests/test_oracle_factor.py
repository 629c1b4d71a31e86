"""Pins the oracle's seed spreading (_SetArgs.py:46-75) and source
factorization (EuclideanFactor.h:8-30, FiniteDifferences.h:30-45) by what
the mathematics fixes:
  * spread cells = the lattice points of the open disk, found by brute force;
    spread values = exact distances for constant metrics, and the exact
    straight-line travel time for a cost linear along the segment;
  * with a constant metric the distance T(x) = |x - seed|_M solves the
    factored scheme exactly (both orders), so the factored fixed point is T
    to rounding for axis stencils, and away from the boundary for Selling
    stencils (whose upwind offsets may leave the grid at the border);
  * the shifts act inside the ball only: the factored solution is exact
    inside it and carries the first-order error outside;
  * on a smooth variable cost (linear speed, closed-form travel time)
    factorization reduces the error near the seed.
"""
import numpy as np
import pytest

M = np.array([[2., 0.6], [0.6, 1.]])
CASES = [("Isotropic", {}), ("Diagonal", {"metric": np.array([1.5, 0.7])}),
         ("Riemann", {"metric": M}), ("Rander", {"metric": M, "drift": np.array([0.3, -0.2])})]


def grid_points(shape):
    return np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing='ij'), -1).reshape(-1, len(shape))


def setup(O, model, kw, n=41, h=(0.05, 0.05), seed=(20., 20.), radius=2.):
    shape = (n, n)
    cm_at = lambda p: O.cost_metric_at(model, p, h, shape, **kw)
    cells, vals = O.spread_seed(np.array(seed), radius, shape, cm_at)
    P = O.make_problem(model, shape, h, seeds_idx=cells, seed_values=vals, **kw)
    cm = cm_at(np.array(seed)[None])
    T = O.metric_norm(cm, grid_points(shape) - np.array(seed)).reshape(shape)
    return P, T, cm, cells, vals


def test_spread_cells_brute_force(O):
    seed = np.array([7.3, 4.6])
    for radius in (1., 2., 2.5):
        cells, _ = O.spread_seed(seed, radius, (10, 10), lambda p: ('diag', np.ones((len(p), 2)), None))
        ref = [(i, j) for i in range(10) for j in range(10) if (i - seed[0]) ** 2 + (j - seed[1]) ** 2 < radius ** 2]
        assert sorted(map(tuple, cells)) == sorted(ref)
    # near a border: cells outside the domain are dropped; periodic axes wrap
    seed = np.array([0.2, 5.])
    disk = [(i, j) for i in range(-3, 4) for j in range(10) if (i - seed[0]) ** 2 + (j - seed[1]) ** 2 < 4.]
    cells, _ = O.spread_seed(seed, 2., (10, 10), lambda p: ('diag', np.ones((len(p), 2)), None))
    assert sorted(map(tuple, cells)) == sorted(c for c in disk if c[0] >= 0)
    cells, _ = O.spread_seed(seed, 2., (10, 10), lambda p: ('diag', np.ones((len(p), 2)), None),
                             periodic=[1, 0])
    assert sorted(map(tuple, cells)) == sorted((i % 10, j) for i, j in disk)


@pytest.mark.parametrize("model,kw", CASES)
def test_spread_values_constant_metric(O, model, kw):
    """Constant metric: both endpoint norms agree, the value is the distance."""
    P, T, cm, cells, vals = setup(O, model, kw, seed=(20.4, 19.7), radius=2.5)
    assert len(cells) >= 16
    np.testing.assert_allclose(vals, T[tuple(cells.T)], rtol=1e-14, atol=1e-15)


def test_spread_values_linear_cost(O):
    """Isotropic cost linear along the segment: the mean of the end costs
    times the length is the exact straight-line travel time."""
    n, h = 30, np.array([0.1, 0.2])
    X = grid_points((n, n)).astype(float)
    cost = (1. + 0.05 * X[:, 0] - 0.03 * X[:, 1]).reshape(n, n)
    seed = np.array([12., 15.])      # on the grid: its cost is exact
    cells, vals = O.spread_seed(seed, 3., (n, n), lambda p: O.cost_metric_at('Isotropic', p, h, (n, n), cost=cost))
    t = np.linspace(0., 1., 2001)
    for c, v in zip(cells, vals):
        d = c - seed
        pts = seed + t[:, None] * d
        ct = 1. + 0.05 * pts[:, 0] - 0.03 * pts[:, 1]
        exact = np.trapezoid(ct, t) * np.linalg.norm(d * h)
        assert abs(v - exact) <= 1e-12 * max(1., exact)


def test_cost_metric_interpolation(O):
    """Order-1 interpolation reproduces multilinear fields exactly."""
    n, h = 12, np.array([0.5, 0.25])
    X = grid_points((n, n)).astype(float)
    c = (2. + 0.1 * X[:, 0] + 0.2 * X[:, 1] + 0.01 * X[:, 0] * X[:, 1]).reshape(n, n)
    p = np.array([[3.3, 7.8], [0., 0.], [10.5, 10.99]])
    kind, costs = O.cost_metric_at('Isotropic', p, h, (n, n), cost=c)[:2]
    ex = 2. + 0.1 * p[:, 0] + 0.2 * p[:, 1] + 0.01 * p[:, 0] * p[:, 1]
    np.testing.assert_allclose(costs, ex[:, None] * h[None], rtol=1e-13)


@pytest.mark.parametrize("model,kw", CASES)
def test_factored_constant_metric_exact(O, model, kw):
    P, T, cm, cells, vals = setup(O, model, kw)
    u0, _ = P.solve_jacobi(1e-13, 0.)
    with O.factor(1e9, (20., 20.), O.factor_metric(cm)):
        u1, _ = P.solve_jacobi(1e-13, 0.)
        # T is a fixed point: the scheme applied to T returns T (interior)
        res = np.abs(P.flow(T)['values'] - T)
        u3, it = P.solve_jacobi(1e-13, 0., u0=T)
    err0, err1 = np.abs(u0 - T), np.abs(u1 - T)
    assert res[4:-4, 4:-4].max() <= 1e-14
    assert it == 0 and np.abs(u3 - T).max() <= 1e-14
    if model in ("Isotropic", "Diagonal"):
        assert err1.max() <= 1e-13                  # axis stencils: exact everywhere
    else:
        assert err1[5:-5, 5:-5].max() <= 1e-6      # border cells may lack an upwind neighbour
        assert err1.min() >= -1e-13                 # ... which only raises values
    assert err0[5:-5, 5:-5].max() >= 1e-2           # the unfactored scheme is first order
    assert err1.max() <= 0.1 * err0.max()


@pytest.mark.parametrize("model,kw", CASES[:3])
def test_factored_order2_fixed_point(O, model, kw):
    """Second order: the further neighbours shifted by fact2 make T a fixed
    point of the corrected scheme as well."""
    P, T, cm, cells, vals = setup(O, model, kw)
    with O.factor(1e9, (20., 20.), O.factor_metric(cm)), O.order(2):
        u, it = P.solve_jacobi(1e-13, 0., u0=T)
        res = np.abs(P.flow(T)['values'] - T)
    assert it == 0 and np.abs(u - T).max() <= 1e-14
    assert res[4:-4, 4:-4].max() <= 1e-14


def test_factor_ball(O):
    """Shifts inside |x - seed| < radius only.  Characteristics leave the
    seed, so inside the ball the factored solution is the distance, while
    outside the first-order error of the unshifted scheme builds up."""
    P, T, cm, cells, vals = setup(O, "Isotropic", {})
    r = 8.
    with O.factor(r, (20., 20.), O.factor_metric(cm)):
        u, _ = P.solve_jacobi(1e-13, 0.)
    X = grid_points((41, 41)).reshape(41, 41, 2)
    dist = np.linalg.norm(X - 20., axis=-1)
    err = np.abs(u - T)
    assert err[dist < r].max() <= 1e-13
    assert err[dist > r + 5].max() >= 1e-3


def test_factor_fast_marching_refused(O):
    P, T, cm, cells, vals = setup(O, "Isotropic", {})
    with O.factor(5., (20., 20.), O.factor_metric(cm)):
        with pytest.raises(ValueError):
            P.solve_fm()


def test_factored_linear_speed_more_accurate(O):
    """Isotropic speed s(x) = s0 + <G, x - x0>: travel time
    arccosh(1 + |G|^2 |x - x0|^2 / (2 s0 s(x))) / |G|.  Factorization
    (metric frozen at the seed) cuts the error near the seed."""
    n = 81
    h = 2. / (n - 1)
    x = (np.arange(n) - n // 2) * h
    X = np.stack(np.meshgrid(x, x, indexing='ij'))
    s0, G = 1., np.array([0.4, 0.25])
    speed = s0 + G[0] * X[0] + G[1] * X[1]
    r2 = X[0] ** 2 + X[1] ** 2
    g = np.linalg.norm(G)
    ex = np.arccosh(1. + g * g * r2 / (2. * s0 * speed)) / g
    seed = np.array([n // 2, n // 2], dtype=float)
    shape = (n, n)
    cm_at = lambda p: O.cost_metric_at('Isotropic', p, [h, h], shape, cost=1. / speed)
    cells, vals = O.spread_seed(seed, 2., shape, cm_at)
    P = O.make_problem('Isotropic', shape, [h, h], cost=1. / speed, seeds_idx=cells, seed_values=vals)
    u0, _ = P.solve_jacobi(1e-13, 0.)
    with O.factor(10., seed, O.factor_metric(cm_at(seed[None]))):
        u1, _ = P.solve_jacobi(1e-13, 0.)
    near = r2 < 0.3 ** 2
    e0, e1 = np.abs(u0 - ex)[near].max(), np.abs(u1 - ex)[near].max()
    assert e1 <= 0.25 * e0, (e0, e1)
