"""The oracle's curvature-penalised models (ReedsShepp2.h, ReedsSheppForward2.h,
Dubins2.h, Elastica2.h with Decomp_v_.h) pinned by what the mathematics fixes:

* the Fejer rule integrates cos^m(phi) sin(phi) over [0, pi] exactly for
  m < 5, and its rounded table equals the reference's (Elastica2.h:40-43);
* decomp_v keeps exactly the offsets of the Selling decomposition of the
  relaxed outer product that lie within acos(sqrt(2/3)) of the direction,
  with their Selling weights, turned against it;
* the discrete Hamiltonians sum_k w_k (-<p, e_k>)_+^2 approach the models'
  continuous ones as the relaxation vanishes;
* closed forms on the grid: straight motion along theta = 0 and rotation in
  place cost exactly h and xi h_theta per cell for ReedsShepp2 and
  ReedsSheppForward2; the forward models cannot back up;
* fast marching and the Jacobi iteration reach the same fixed point.
"""
import numpy as np
import pytest

CURV = ['ReedsShepp2', 'ReedsSheppForward2', 'Dubins2', 'Elastica2']


def cell_problem(O, model, ixi, kappa, theta):
    """One independent cell per geometry entry (shape (n, 1, 1))."""
    n = len(ixi)
    geom = np.stack([ixi, kappa, np.cos(theta), np.sin(theta)])
    return O.Problem(model, (n, 1, 1), geom, np.ones(n), np.zeros(n, np.uint8), [0., 0., 0.])


def relaxed(v, eps2):
    """(1 - eps^2) v v^T + eps^2 |v|^2 Id: eigenvalue |v|^2 along v and
    eps^2 |v|^2 across it (Geometry_.h:199-210)."""
    return (1 - eps2) * np.outer(v, v) + eps2 * np.dot(v, v) * np.eye(3)


def flat(m):
    return np.array([m[0, 0], m[1, 0], m[1, 1], m[2, 0], m[2, 1], m[2, 2]])


def test_fejer_rule(O):
    t = O.fejer(5)
    phi = np.pi * (np.arange(5) + 0.5) / 5
    np.testing.assert_allclose(t[:, 0], np.cos(phi), atol=1e-9)
    np.testing.assert_allclose(t[:, 1], np.sin(phi), atol=1e-9)
    assert t[2, 0] == 0.                    # cos(pi/2) exactly
    for m in range(5):
        exact = (1 + (-1) ** m) / (m + 1)   # int_0^pi cos^m sin
        assert abs(np.sum(t[:, 2] * t[:, 0] ** m) - exact) <= 3e-6
    # the reference's 6-decimal table (Elastica2.h:40-43)
    np.testing.assert_array_equal(t[:, 2], [0.167781, 0.525552, 0.613333, 0.525552, 0.167781])


def decomp_check(O, v, w, e, eps2=0.01):
    """decomp_v(v) -> (w[6], e[6, 3]) against the Selling decomposition of
    the relaxed outer product."""
    ws, es = O.selling(flat(relaxed(v, eps2))[None])
    ws, es = ws[0], es[0]
    vv = v @ v
    kept = {}
    for k in range(6):
        ve = v @ es[k]
        if ve * ve >= vv * (es[k] @ es[k]) * (2. / 3.) and ws[k] > 0:
            kept[tuple(-es[k] if ve > 0 else es[k])] = ws[k]
    got = {tuple(e[k]): w[k] for k in range(6) if w[k] > 0}
    assert got.keys() == kept.keys()
    for key in got:
        assert got[key] == pytest.approx(kept[key], rel=1e-12)
        assert np.dot(v, key) <= 0            # the forward neighbour x + e lies upwind


def test_decomp_v_is_pruned_selling(O):
    rng = np.random.default_rng(3)
    n = 200
    ixi, kappa, th = rng.uniform(0.2, 3., n), rng.uniform(-1., 1., n), rng.uniform(0, 2 * np.pi, n)
    P = cell_problem(O, 'ReedsSheppForward2', ixi, kappa, th)
    w, e = P.stencils()
    for c in range(n):
        assert w[c, 0] == pytest.approx(ixi[c] ** 2, rel=1e-15)
        assert tuple(e[c, 0]) == (0, 0, -1)
        decomp_check(O, np.array([np.cos(th[c]), np.sin(th[c]), kappa[c]]), w[c, 1:], e[c, 1:])
    P = cell_problem(O, 'Dubins2', ixi, kappa, th)
    w, e = P.stencils()
    for c in range(0, n, 7):
        for m, sg in ((0, 1), (1, -1)):
            v = np.array([np.cos(th[c]), np.sin(th[c]), kappa[c] + sg * ixi[c]])
            decomp_check(O, v, w[c, 6 * m:6 * m + 6], e[c, 6 * m:6 * m + 6])


def forward_hamiltonian(w, e, p):
    return np.sum(w * np.maximum(0., -(e @ p)) ** 2)


@pytest.mark.parametrize("eps", [0.2, 0.1, 0.05])
def test_forward_hamiltonians_are_consistent(O, eps):
    """sum_k w_k (-<p, e_k>)_+^2 against (<p, v>)_+^2 (ReedsSheppForward2's
    forward part and each Dubins2 mix), and against
    sum_l w_l (<p, v_l>)_+^2 over the Fejer nodes (Elastica2): the error
    scales with the relaxation eps^2 (1.2 to 1.9 eps^2 at worst over these
    samples), far below that of a sign or threshold mistake.  (Much smaller
    relaxations leave matrices too flat for the Selling iteration cap.)"""
    rng = np.random.default_rng(5)
    n = 300
    ixi, kappa, th = rng.uniform(0.3, 2., n), rng.uniform(-0.5, 0.5, n), rng.uniform(0, 2 * np.pi, n)
    p = rng.normal(size=(n, 3))
    fej = O.fejer(5)
    with O.curvature_relax(eps * eps):
        rsf = cell_problem(O, 'ReedsSheppForward2', ixi, kappa, th).stencils()
        dub = cell_problem(O, 'Dubins2', ixi, kappa, th).stencils()
        ela = cell_problem(O, 'Elastica2', ixi, kappa, th).stencils()
    worst = 0.
    for c in range(n):
        v = np.array([np.cos(th[c]), np.sin(th[c]), kappa[c]])
        scale = (p[c] @ p[c]) * (v @ v)
        h_d = forward_hamiltonian(rsf[0][c, 1:], rsf[1][c, 1:], p[c])
        worst = max(worst, abs(h_d - max(0., p[c] @ v) ** 2) / scale)
        for m, sg in ((0, 1), (1, -1)):
            vm = v + np.array([0, 0, sg * ixi[c]])
            h_d = forward_hamiltonian(dub[0][c, 6 * m:6 * m + 6], dub[1][c, 6 * m:6 * m + 6], p[c])
            worst = max(worst, abs(h_d - max(0., p[c] @ vm) ** 2) / ((p[c] @ p[c]) * (vm @ vm)))
        ref = 0.
        for cP, sP, wF in fej:
            vl = np.array([sP * v[0], sP * v[1], sP * kappa[c] + cP * ixi[c]])
            ref += wF * max(0., p[c] @ vl) ** 2
        h_d = forward_hamiltonian(ela[0][c], ela[1][c], p[c])
        worst = max(worst, abs(h_d - ref) / ((p[c] @ p[c]) * (v @ v + ixi[c] ** 2)))
    assert worst <= 2 * eps * eps, worst


def test_reeds_shepp_hamiltonian(O):
    """ReedsShepp2: sum_k w_k <p, e_k>^2 <= p^T M p, M the relaxed outer
    product with its angular entry raised to kappa^2 + 1/xi^2 (pruning drops
    positive terms only), and the dropped part is of order eps^2."""
    rng = np.random.default_rng(6)
    n = 300
    ixi, kappa, th = rng.uniform(0.3, 2., n), rng.uniform(-0.5, 0.5, n), rng.uniform(0, 2 * np.pi, n)
    w, e = cell_problem(O, 'ReedsShepp2', ixi, kappa, th).stencils()
    p = rng.normal(size=(n, 3))
    for c in range(n):
        v = np.array([np.cos(th[c]), np.sin(th[c]), kappa[c]])
        M = relaxed(v, 0.01)
        M[2, 2] = max(M[2, 2], kappa[c] ** 2 + ixi[c] ** 2)
        h_d = np.sum(w[c] * (e[c] @ p[c]) ** 2)
        full = p[c] @ M @ p[c]
        assert h_d <= full * (1 + 1e-12)
        assert full - h_d <= 0.05 * (p[c] @ p[c]) * (v @ v)


def test_axis_aligned_stencils(O):
    """theta = 0, kappa = 0: the relaxed matrix is diagonal; ReedsShepp2 keeps
    e_x (weight 1) and e_theta (weight 1/xi^2), ReedsSheppForward2 the
    rotation (weight 1/xi^2) and the backward neighbour -e_x (weight 1)."""
    ixi = np.array([1.7, 0.4])
    z = np.zeros(2)
    w, e = cell_problem(O, 'ReedsShepp2', ixi, z, z).stencils()
    for c in range(2):
        got = {tuple(e[c, k]): w[c, k] for k in range(6) if w[c, k] > 0}
        assert got == {(1, 0, 0): 1.0, (0, 0, 1): pytest.approx(max(0.01, ixi[c] ** 2))}
    w, e = cell_problem(O, 'ReedsSheppForward2', ixi, z, z).stencils()
    for c in range(2):
        got = {tuple(e[c, k]): w[c, k] for k in range(7) if w[c, k] > 0}
        assert got == {(0, 0, -1): pytest.approx(ixi[c] ** 2), (-1, 0, 0): 1.0}


def grid_problem(O, model, n=21, nt=32, xi=0.3):
    h = np.array([2. / n, 2. / n, 2 * np.pi / nt])
    geom = O.curvature_geometry((n, n, nt), h, xi)
    rhs = np.full((n, n, nt), h[0])               # unit cost times h
    seeds = np.zeros((n, n, nt), np.uint8)
    c = n // 2
    seeds[c, c, 0] = 1
    rhs[c, c, 0] = 0.
    return O.Problem(model, (n, n, nt), geom, rhs, seeds, [0., 0., 0.], periodic=[0, 0, 1]), h, c


@pytest.mark.parametrize("model", ['ReedsShepp2', 'ReedsSheppForward2'])
def test_straight_motion_and_rotation_closed_forms(O, model):
    xi = 0.3
    P, h, c = grid_problem(O, model, xi=xi)
    u = P.solve_fm()
    k = np.arange(1, 8)
    np.testing.assert_allclose(u[c + k, c, 0], k * h[0], rtol=1e-13)          # ahead
    np.testing.assert_allclose(u[c, c, k], k * h[2] * xi, rtol=1e-13)         # turning in place
    np.testing.assert_allclose(u[c, c, -k], k * h[2] * xi, rtol=1e-13)        # periodic angle
    behind = u[c - k, c, 0] / (k * h[0])
    if model == 'ReedsShepp2':
        np.testing.assert_allclose(behind, 1., rtol=1e-13)                    # reverse gear
    else:
        # no reverse gear: turn around in place twice (cost 2 pi xi), then go
        extra = u[c - k, c, 0] - k * h[0]
        assert np.all(extra >= 0.95 * 2 * np.pi * xi) and np.all(extra <= 1.1 * 2 * np.pi * xi)


@pytest.mark.parametrize("model", ['Dubins2', 'Elastica2'])
def test_forward_models_cannot_back_up(O, model):
    P, h, c = grid_problem(O, model)
    u = P.solve_fm()
    k = np.arange(2, 8)
    assert np.all(u[c - k, c, 0] > 2.5 * u[c + k, c, 0])


@pytest.mark.parametrize("model", CURV)
def test_fast_marching_equals_iteration(O, model):
    P, h, c = grid_problem(O, model, n=15, nt=16, xi=0.5)
    u = P.solve_fm()
    uj, it = P.solve_jacobi(1e-13, 0.)
    fin = np.isfinite(u)
    assert np.array_equal(fin, np.isfinite(uj))
    assert np.max(np.abs(u[fin] - uj[fin])) <= 1e-12 * np.max(u[fin])
    # (Dubins2 turns at a bounded curvature: states near the box edges that
    # would need a tighter turn stay unreachable)
    assert np.all(u[fin] >= 0) and fin.sum() >= (0.75 if model == 'Dubins2' else 1.) * u.size


@pytest.mark.parametrize("model", CURV)
def test_flow_weights_of_forward_offsets(O, model):
    """The flow of each cell points from the upwind neighbours: a forward
    offset contributes only on its x + e side."""
    P, h, c = grid_problem(O, model, n=15, nt=16, xi=0.5)
    u = P.solve_fm()
    fl = P.flow(u)
    fin = np.isfinite(u)
    assert np.array_equal(np.isfinite(fl['values']), fin)
    assert np.max(np.abs(fl['values'][fin] - u[fin])) <= 1e-12 * np.max(u[fin])
    assert np.all(fl['weights'] >= 0)
    ns, N = P.nsym, P.N
    nfwd = {'ReedsShepp2': 0, 'ReedsSheppForward2': 6, 'Dubins2': 6, 'Elastica2': 30}[model]
    w, e = P.stencils()
    shape = np.array(P.shape)
    cells = np.stack(np.unravel_index(np.arange(N), P.shape), -1)
    mixed = P.nmix > 1
    for k in range(ns - nfwd, ns):
        live = fl['indices'][:, k] >= 0
        if mixed:
            continue        # the chosen mix is not exported; checked through the values
        y = cells[live] + e[live, k]
        y[:, 2] %= shape[2]
        np.testing.assert_array_equal(np.ravel_multi_index(tuple(y.T), P.shape), fl['indices'][live, k])


def test_convex_variants(O):
    """Convex variants (left turns only, kernel_traits.py:129-130): the
    rotation of ReedsSheppForward2 turns left only, so turning right by k
    cells costs (n_theta - k) xi h_theta; the right-turn mix of Dubins2 is
    straight motion and the Elastica2 nodes weigh 1, 1/2, 0 (Hamiltonians as
    above); fast marching equals the iteration."""
    xi = 0.3
    with O.convex():
        P, h, c = grid_problem(O, 'ReedsSheppForward2', xi=xi)
        u = P.solve_fm()
        nt = P.shape[2]
        k = np.arange(1, 8)
        np.testing.assert_allclose(u[c, c, k], k * h[2] * xi, rtol=1e-13)
        np.testing.assert_allclose(u[c, c, nt - k], (nt - k) * h[2] * xi, rtol=1e-13)
        np.testing.assert_allclose(u[c + k, c, 0], k * h[0], rtol=1e-13)
        rng = np.random.default_rng(8)
        n = 200
        ixi, kappa, th = rng.uniform(0.3, 2., n), rng.uniform(-0.5, 0.5, n), rng.uniform(0, 2 * np.pi, n)
        dub = cell_problem(O, 'Dubins2', ixi, kappa, th).stencils()
        ela = cell_problem(O, 'Elastica2', ixi, kappa, th).stencils()
        for model in ('Dubins2', 'Elastica2', 'ReedsSheppForward2'):
            Pm, _, _ = grid_problem(O, model, n=15, nt=16, xi=0.5)
            a = Pm.solve_fm()
            b, _ = Pm.solve_jacobi(1e-13, 0.)
            fin = np.isfinite(a)
            assert np.array_equal(fin, np.isfinite(b))
            assert np.max(np.abs(a[fin] - b[fin])) <= 1e-12 * np.max(a[fin])
    p = rng.normal(size=(n, 3))
    fej = O.fejer(5)
    factor = [1., 1., 0.5, 0., 0.]
    worst = 0.
    for cc in range(n):
        v = np.array([np.cos(th[cc]), np.sin(th[cc]), kappa[cc]])
        vm = v.copy()                         # the second Dubins mix: straight ahead
        h_d = forward_hamiltonian(dub[0][cc, 6:], dub[1][cc, 6:], p[cc])
        worst = max(worst, abs(h_d - max(0., p[cc] @ vm) ** 2) / ((p[cc] @ p[cc]) * (vm @ vm)))
        ref = 0.
        for (cP, sP, wF), f in zip(fej, factor):
            vl = np.array([sP * v[0], sP * v[1], sP * kappa[cc] + cP * ixi[cc]])
            ref += f * wF * max(0., p[cc] @ vl) ** 2
        h_d = forward_hamiltonian(ela[0][cc], ela[1][cc], p[cc])
        worst = max(worst, abs(h_d - ref) / ((p[cc] @ p[cc]) * (v @ v + ixi[cc] ** 2)))
    assert worst <= 2 * 0.01, worst
    # the Elastica2 stencil of the right-turn nodes is empty
    assert np.all(ela[0][:, 18:] == 0.)
