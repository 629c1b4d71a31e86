"""Pins the oracle's solvers, flow and adjoint with what the mathematics fixes:
schedule independence of the fixed point, the scheme residual, closed-form
distances, finite differences of the FP64 solve, and the reference's stored
known answer for the seed-value sensitivity."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def small_riemann(O, n=24, seed=0, d=2, amax=4.):
    rng = np.random.default_rng(seed)
    shape = (n,) * d
    A = rng.normal(size=(d, d))
    # smooth anisotropic field: rotate a fixed diagonal metric by a smooth angle
    if d == 2:
        x = np.linspace(0, 1, n)
        th = np.add.outer(2 * x, -x)
        c, s = np.cos(th), np.sin(th)
        m = np.array([[c * c + amax * s * s, (amax - 1) * c * s], [(amax - 1) * c * s, s * s + amax * c * c]])
    else:
        M = A @ A.T + np.eye(d)
        m = np.broadcast_to(M.reshape(d, d, *(1,) * d), (d, d) + shape).copy()
    h = np.full(d, 1. / n)
    seeds = [[n // 2] * d]
    return O.make_problem('Riemann', shape, h, metric=m, seeds_idx=seeds), m, h


@pytest.mark.parametrize("model", ["Isotropic", "Riemann"])
@pytest.mark.parametrize("d", [2, 3])
def test_fast_marching_equals_jacobi_fixed_point(O, model, d):
    n = 20 if d == 2 else 9
    if model == "Riemann":
        P, _, _ = small_riemann(O, n, d=d)
    else:
        rng = np.random.default_rng(1)
        P = O.make_problem('Isotropic', (n,) * d, [1. / n] * d, cost=1 + 0.5 * rng.uniform(size=(n,) * d),
                           seeds_idx=[[1] * d])
    u_fm = P.solve_fm()
    u_j, it = P.solve_jacobi(1e-14, 1e-14)
    assert it < 100000
    np.testing.assert_allclose(u_j, u_fm, rtol=1e-10, atol=1e-12)


def test_scheme_residual_of_fast_marching(O):
    """The converged values satisfy their own update (one more flow pass
    changes nothing) -- the reference's residual check pattern."""
    P, _, _ = small_riemann(O, 32)
    u = P.solve_fm()
    f = P.flow(u)
    np.testing.assert_allclose(f['values'], u, rtol=1e-13, atol=1e-13)


def test_isotropic_distance_converges(O):
    """Constant isotropic cost: u -> |x - s| at rate O(h)."""
    errs = []
    for n in (17, 33, 65):
        h = 2. / (n - 1)
        P = O.make_problem('Isotropic', (n, n), [h, h], cost=1., seeds_idx=[[n // 2, n // 2]])
        u = P.solve_fm()
        x = (np.arange(n) - n // 2) * h
        ex = np.sqrt(np.add.outer(x ** 2, x ** 2))
        errs.append(np.max(np.abs(u - ex)))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.08


def test_riemann_constant_metric_distance(O):
    """Constant metric M: u -> sqrt(<x-s, M (x-s)>) with refinement."""
    M = np.array([[2., 0.6], [0.6, 1.]])
    errs = []
    for n in (31, 61, 121):
        h = 2. / (n - 1)
        P = O.make_problem('Riemann', (n, n), [h, h], metric=M, seeds_idx=[[n // 2, n // 2]])
        u = P.solve_fm()
        x = (np.arange(n) - n // 2) * h
        X = np.stack(np.meshgrid(x, x, indexing='ij'))
        ex = np.sqrt(np.einsum('i...,ij,j...->...', X, M, X))
        errs.append(np.max(np.abs(u - ex)))
        assert np.all(u >= ex - 1e-12)  # the upwind scheme over-estimates the distance
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.05


def test_flow_points_away_from_seed(O):
    n = 41
    h = 2. / (n - 1)
    P = O.make_problem('Riemann', (n, n), [h, h], metric=np.eye(2), seeds_idx=[[n // 2, n // 2]])
    u = P.solve_fm()
    f = P.flow(u)
    flow = -h * f['vector']  # hfmOut['flow'] convention (_PostProcess.py:89-91)
    x = (np.arange(n) - n // 2) * h
    X = np.stack(np.meshgrid(x, x, indexing='ij'))
    r = np.linalg.norm(X, axis=0)
    mask = r > 0.3
    cos = np.sum(flow * X, axis=0)[mask] / (np.linalg.norm(flow, axis=0)[mask] * r[mask])
    assert np.min(cos) > 0.9


def test_adjoint_matches_finite_differences(O):
    """d(sum of receiver values)/d cost(x) = cost(x) * lambda(x), with lambda
    the exact reverse solve (costSensitivity convention of the reference)."""
    n = 20
    rng = np.random.default_rng(3)
    cost = 1 + 0.3 * rng.uniform(size=(n, n))
    h = 1. / n
    th = rng.uniform(0, np.pi)
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    M = R.T @ np.diag([1., 3.]) @ R
    seeds = [[3, 4]]
    receivers = [(n - 1, n - 2), (n - 3, 1), (2, n - 1)]

    def J(c):
        P = O.make_problem('Riemann', (n, n), [h, h], metric=M, cost=c, seeds_idx=seeds)
        u = P.solve_fm()
        return sum(u[r] for r in receivers), P, u

    J0, P, u = J(cost)
    sens = np.zeros((n, n, 1))
    for r in receivers:
        sens[r + (0,)] = 1.
    lam = P.adjoint(u, P.flow(u), sens)[..., 0]
    grad = cost * lam
    cells = [tuple(c) for c in rng.integers(0, n, (12, 2))]
    eps = 1e-6
    for c in cells:
        if c == tuple(seeds[0]):
            continue
        cp, cm = cost.copy(), cost.copy()
        cp[c] += eps
        cm[c] -= eps
        fd = (J(cp)[0] - J(cm)[0]) / (2 * eps)
        assert fd == pytest.approx(grad[c], rel=1e-4, abs=1e-7)


def test_adjoint_identity(O):
    """<s, du> = <lambda, c dc> for a random perturbation (first order)."""
    n = 16
    rng = np.random.default_rng(4)
    cost = 1 + 0.5 * rng.uniform(size=(n, n))
    h = 1. / n
    P = O.make_problem('Isotropic', (n, n), [h, h], cost=cost, seeds_idx=[[2, 3]])
    u = P.solve_fm()
    s = rng.uniform(size=(n, n, 1))
    lam = P.adjoint(u, P.flow(u), s)[..., 0]
    dc = rng.normal(size=(n, n)) * 1e-7
    dc[2, 3] = 0
    P2 = O.make_problem('Isotropic', (n, n), [h, h], cost=cost + dc, seeds_idx=[[2, 3]])
    du = P2.solve_fm() - u
    assert np.sum(s[..., 0] * du) == pytest.approx(np.sum(lam * cost * dc), rel=1e-4)


def test_seed_value_sensitivity_kat(O):
    k = json.load(open(os.path.join(HERE, "golden", "ad_kat.json")))["seed_value_sensitivity"]
    n = k['dimx']
    h = 1. / (n - 1)  # sampleBoundary=True: cell centres on the box corners
    seeds = np.round((np.asarray(k['seeds']) - 0.) / h).astype(int)
    P = O.make_problem('Isotropic', (n, n), [h, h], cost=k['cost'], seeds_idx=seeds)
    u = P.solve_fm()
    sens = np.zeros((n, n, 1))
    sens[tuple(k['sensitivity_cell']) + (0,)] = 1.
    lam = P.adjoint(u, P.flow(u), sens)
    got = np.array([lam[tuple(s)] for s in seeds])
    np.testing.assert_allclose(got, k['expected'], atol=k['tol'])


def test_walls_distance_matches_bfs(O):
    rng = np.random.default_rng(2)
    walls = rng.uniform(size=(23, 17)) < 0.1
    dist = O.wall_distance(walls, bound=10)
    # brute force: L1 distance to the nearest wall, capped
    wi = np.argwhere(walls)
    X = np.stack(np.meshgrid(np.arange(23), np.arange(17), indexing='ij'), -1)
    l1 = np.min(np.abs(X[:, :, None, :] - wi[None, None]).sum(-1), axis=-1)
    ref = np.where(l1 > 10, 255, l1)
    assert np.array_equal(dist, ref)


def test_walls_block_the_front(O):
    """A wall across the domain with one gap: cells behind it are reached
    through the gap only."""
    n = 31
    h = 1. / n
    walls = np.zeros((n, n), dtype=bool)
    walls[15, :] = True
    walls[15, 25] = False
    P = O.make_problem('Isotropic', (n, n), [h, h], cost=1., seeds_idx=[[5, 5]], walls=walls)
    u = P.solve_fm()
    assert np.all(np.isinf(u[walls]))
    # behind the wall, right below the seed: path must detour via column 25
    direct = 20 * h
    assert u[25, 5] > direct + 10 * h


def test_forward_ad_matches_finite_differences(O):
    """du along a cost and seed-value perturbation = central FD of the FP64
    solve (rhs = cost * dcost off the seeds, dseed on the seeds)."""
    n = 20
    rng = np.random.default_rng(5)
    cost = 1 + 0.3 * rng.uniform(size=(n, n))
    h = 1. / n
    M = np.array([[2., 0.5], [0.5, 1.]])
    seeds = [[3, 4], [15, 12]]
    sv = np.array([0., 0.2])
    dc = rng.normal(size=(n, n))
    dsv = np.array([0.3, -0.5])

    def solve(c, s):
        P = O.make_problem('Riemann', (n, n), [h, h], metric=M, cost=c, seeds_idx=seeds, seed_values=s)
        return P.solve_fm(), P

    u, P = solve(cost, sv)
    r = cost * dc
    for (i, j), d in zip(seeds, dsv):
        r[i, j] = d
    du = P.forward_ad(u, P.flow(u), r[..., None])[..., 0]
    eps = 1e-6
    fd = (solve(cost + eps * dc, sv + eps * dsv)[0] - solve(cost - eps * dc, sv - eps * dsv)[0]) / (2 * eps)
    np.testing.assert_allclose(du, fd, rtol=1e-4, atol=1e-7)


def test_forward_ad_homogeneity(O):
    """u(l c, l s) = l u(c, s), so the variation along (c, s) is u itself
    (the sum rule of the seed/cost sensitivities)."""
    n = 18
    rng = np.random.default_rng(6)
    cost = 1 + 0.5 * rng.uniform(size=(n, n, n // 2))
    P = O.make_problem('Riemann', (n, n, n // 2), [1. / n] * 3, metric=np.diag([1., 2., 3.]), cost=cost,
                       seeds_idx=[[2, 3, 4]], seed_values=[0.25])
    u = P.solve_fm()
    r = cost * cost
    r[2, 3, 4] = 0.25
    du = P.forward_ad(u, P.flow(u), r[..., None])[..., 0]
    np.testing.assert_allclose(du, u, rtol=1e-10)


def test_forward_and_reverse_are_transposes(O):
    """<s, du> computed forward equals <lambda, r> computed in reverse."""
    n = 16
    rng = np.random.default_rng(8)
    cost = 1 + 0.5 * rng.uniform(size=(n, n))
    P = O.make_problem('Isotropic', (n, n), [1. / n] * 2, cost=cost, seeds_idx=[[4, 9]])
    u = P.solve_fm()
    fl = P.flow(u)
    r = rng.normal(size=(n, n, 1))
    s = rng.normal(size=(n, n, 1))
    du = P.forward_ad(u, fl, r)
    lam = P.adjoint(u, fl, s)
    assert np.sum(s * du) == pytest.approx(np.sum(lam * r), rel=1e-12)


def _round_half_away(n, d):
    """n / d rounded to the nearest integer, halves away from zero (d > 0)."""
    q = (2 * abs(n) + d) // (2 * d)
    return q if n >= 0 else -q


def _visible_brute(walls, x, e, periodic):
    """Definition of visibility: no sample x + round(e t / |e|_inf),
    t = 1 .. |e|_inf, lies on a wall (samples outside a non-periodic grid are
    free)."""
    linf = max(abs(c) for c in e)
    shape = walls.shape
    for t in range(1, linf + 1):
        y = []
        inside = True
        for k, (xk, ek) in enumerate(zip(x, e)):
            yk = xk + _round_half_away(ek * t, linf)
            if periodic[k]:
                yk %= shape[k]
            elif not 0 <= yk < shape[k]:
                inside = False
            y.append(yk)
        if inside and walls[tuple(y)]:
            return False
    return True


@pytest.mark.parametrize("d,periodic", [(2, (0, 0)), (2, (1, 0)), (3, (0, 0, 0)), (3, (0, 1, 1))])
def test_wall_visibility_walk_matches_its_definition(O, d, periodic):
    """The oracle's visibility walk (Walls.h:18-50), which skips ahead using
    the saturated wall distance, against the plain definition on random wall
    masks, for every primitive offset with |e|_inf <= 5 (the Selling offsets
    of the anisotropic test metrics are among them).  With the wall distance
    computed up to a bound above every |e|_1, the walk's early exits are exact;
    with the default bound 10 it is checked on the offsets of |e|_1 <= 10."""
    import itertools
    import math
    rng = np.random.default_rng(7 + d)
    shape = (23, 19) if d == 2 else (11, 9, 10)
    walls = rng.random(shape) < (0.12 if d == 2 else 0.08)
    r = 5
    offsets = [e for e in itertools.product(range(-r, r + 1), repeat=d)
               if any(e) and math.gcd(*[abs(c) for c in e]) == 1]
    cells = [x for x in itertools.product(*[range(s) for s in shape]) if not walls[x]]
    cells = [cells[i] for i in rng.choice(len(cells), size=min(len(cells), 60), replace=False)]
    for bound, emax in ((16, 15), (10, 10)):
        dist = O.wall_distance(walls.astype(np.uint8), bound=bound, periodic=list(periodic))
        xs, es = [], []
        for x in cells:
            for e in offsets:
                if sum(abs(c) for c in e) <= emax:
                    xs.append(x)
                    es.append(e)
        got = O.visible(dist, np.array(xs), np.array(es), periodic=list(periodic))
        ref = np.array([_visible_brute(walls, x, e, periodic) for x, e in zip(xs, es)])
        assert (got == ref).all(), [(xs[i], es[i]) for i in np.flatnonzero(got != ref)[:5]]
        # both outcomes occur: the check is not vacuous
        assert 0.05 < ref.mean() < 0.95


def test_periodic_axis_equals_replicated_seeds(O):
    """Periodic wrap of the oracle's grid (Grid.h:68-80): on a grid periodic
    along axis 0 with period n, the fast-marching solution equals the middle
    third of the non-periodic solution on 3n cells with the seeds replicated
    at s, s + n, s + 2n (every upwind chain of a middle cell stays within half
    a period of it, away from the enlarged grid's border).  Riemannian
    anisotropic metric, so that long Selling offsets cross the wrap."""
    rng = np.random.default_rng(3)
    n, m = 24, 17
    h = np.array([1. / n, 1. / m])
    A = rng.normal(size=(2, 2))
    M = A @ A.T + 0.3 * np.eye(2)
    metric = np.broadcast_to(M[:, :, None, None], (2, 2, n, m)).copy()
    metric *= (1. + 0.3 * np.sin(2 * np.pi * np.arange(n) / n))[None, None, :, None]
    seeds = [(3, 5), (17, 12)]
    P = O.make_problem('Riemann', (n, m), h, metric=metric, seeds_idx=np.array(seeds),
                       periodic=[1, 0])
    u = P.solve_fm()
    big = np.concatenate([metric] * 3, axis=2)
    seeds3 = np.array([(s0 + k * n, s1) for k in range(3) for s0, s1 in seeds])
    P3 = O.make_problem('Riemann', (3 * n, m), h, metric=big, seeds_idx=seeds3)
    u3 = P3.solve_fm()
    np.testing.assert_allclose(u, u3[n:2 * n], rtol=0, atol=1e-12)
    # and the wrap is used: the non-periodic solution differs
    Pn = O.make_problem('Riemann', (n, m), h, metric=metric, seeds_idx=np.array(seeds))
    assert np.max(np.abs(Pn.solve_fm() - u)) > 1e-2


def test_periodic_isotropic_torus_distance(O):
    """Constant isotropic metric on a grid periodic along both axes: the
    solution converges to the torus distance, the minimum over the seed's
    periodic images of |x - s - k L|, at the O(h log 1/h) rate of a point
    source (errors 0.030, 0.018, 0.011 for n = 32, 64, 128), while the
    distance without the images stays O(1) away."""
    errs = []
    for n in (32, 64, 128):
        h = 1. / n
        s = (n // 4, n // 3)
        P = O.make_problem('Isotropic', (n, n), (h, h), seeds_idx=np.array([s]), periodic=[1, 1])
        u = P.solve_fm()
        x = (np.arange(n) + 0.5) * h
        X, Y = np.meshgrid(x, x, indexing='ij')
        sx, sy = (s[0] + 0.5) * h, (s[1] + 0.5) * h
        dx = np.abs(X - sx)
        dy = np.abs(Y - sy)
        exact = np.hypot(np.minimum(dx, 1 - dx), np.minimum(dy, 1 - dy))
        errs.append(np.max(np.abs(u - exact)))
        assert np.max(np.abs(np.hypot(dx, dy) - exact)) > 0.4
    assert errs[1] < 0.7 * errs[0] and errs[2] < 0.7 * errs[1]
    assert errs[2] < 0.012
