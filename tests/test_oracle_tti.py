"""Pins the oracle's TTI scheme (TTI_.h): the tangent-ellipse sampling
against the closed-form curve, the reduction to a Riemannian metric when the
quadratic part vanishes, and convergence to the exact TTI distance of a
constant medium (which fixes the min/max choice of the 7 ellipses)."""
import numpy as np
import pytest

HEX = dict(c11=15., c13=4., c33=10., c44=3.)


def hexagonal(c11, c13, c33, c44):
    """TTI.from_hexagonal (Metrics/Seismic/tti.py:141-146)."""
    linear = np.array([c11 + c44, c33 + c44])
    mixed = c13 ** 2 - c11 * c33 + 2. * c13 * c44
    quadratic = np.array([[-2. * c11 * c44, mixed], [mixed, -2. * c33 * c44]])
    return linear, quadratic


def dual_radius(l, q, c):
    """Radius R = r^2 of the dual unit ball along direction c = (cos^2, sin^2):
    smallest positive root of 1/2 <c,qc> R^2 + <l,c> R - 1 = 0."""
    A = 0.5 * np.einsum('i...,ij,j...->...', c, q, c)
    B = np.einsum('i,i...->...', l, c)
    disc = np.sqrt(B * B + 4 * A)
    # (B - disc) / (-2A), written without cancellation: 2 / (B + disc)
    return 2. / (B + disc)


def tti_norm(l, q, A, x):
    """N(x) = max over the dual unit ball of <p, A x> (2D), sampled finely."""
    th = np.linspace(0, 2 * np.pi, 20001)
    c = np.stack([np.cos(th) ** 2, np.sin(th) ** 2])
    r = np.sqrt(dual_radius(l, q, c))
    P = r * np.stack([np.cos(th), np.sin(th)])          # boundary of the dual ball
    y = np.einsum('ij,j...->i...', A, x)
    return np.max(np.einsum('ik,i...->k...', P, y[:, None]), axis=0)


def test_tti_diags_tangent_to_the_dual_curve(O):
    l, qm = hexagonal(**HEX)
    q = np.array([qm[0, 0], qm[1, 0], qm[1, 1]])
    diag, mix = O.tti_diags(l, q)
    diag = diag[0]
    # endpoint closed forms: x = (a, 0) with q00 a^2 + 2 l0 a - 2 = 0
    a = (-l[0] + np.sqrt(l[0] ** 2 + 2 * qm[0, 0])) / qm[0, 0]
    g = np.array([l[0] + qm[0, 0] * a, l[1] + qm[1, 0] * a]) / (a * (l[0] + qm[0, 0] * a))
    np.testing.assert_allclose(diag[0], g, rtol=1e-12)
    # each line <diag_i, y> = 1 touches the curve f(y) = 0 (y >= 0) and leaves
    # it on one side; all lines on the same side
    th = np.linspace(0, np.pi / 2, 4001)
    c = np.stack([np.cos(th) ** 2, np.sin(th) ** 2])
    Y = dual_radius(l, qm, c) * c                       # curve points (X^2, Z^2)
    sides = []
    for dv in diag:
        s = dv @ Y - 1.
        assert abs(np.min(np.abs(s))) < 1e-3            # touches the curve
        sides.append(np.sign(np.median(s)))
        assert np.all(s * sides[-1] >= -1e-6)           # one side only
    assert len(set(sides)) == 1
    # determinant rule of diags(): mix_is_min
    assert bool(mix[0]) == (diag[0, 0] * diag[-1, 1] - diag[0, 1] * diag[-1, 0] > 0)


def test_tti_without_quadratic_part_is_riemannian(O):
    """q = 0: the dual ball is the ellipse l0 X^2 + l1 Z^2 <= 1 of the
    transformed vector, i.e. the Riemannian metric A^T diag(1/l) A."""
    n = 33
    h = np.array([1. / n, 1.3 / n])
    l = np.array([2., 0.7])
    th = 0.4
    A = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]) @ np.diag([1., 1.5])
    P_tti = O.make_problem('TTI', (n, n), h, metric=(l, np.zeros((2, 2)), A), seeds_idx=[[10, 20]])
    M = A.T @ np.diag(1. / l) @ A
    P_r = O.make_problem('Riemann', (n, n), h, metric=M, seeds_idx=[[10, 20]])
    np.testing.assert_allclose(P_tti.solve_fm(), P_r.solve_fm(), rtol=1e-10)


@pytest.mark.parametrize("tilt", [0., 0.5])
def test_tti_converges_to_exact_distance(O, tilt):
    l, qm = hexagonal(**HEX)
    A = np.array([[np.cos(tilt), -np.sin(tilt)], [np.sin(tilt), np.cos(tilt)]])
    errs = []
    for n in (31, 61, 121):
        hh = 2. / (n - 1)
        s = n // 2
        P = O.make_problem('TTI', (n, n), [hh, hh], metric=(l, qm, A), seeds_idx=[[s, s]])
        u = P.solve_fm()
        x = (np.arange(n) - s) * hh
        X = np.stack(np.meshgrid(x, x, indexing='ij'))
        ex = tti_norm(l, qm, A, X)
        errs.append(np.max(np.abs(u - ex)) / np.max(ex))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.03


def test_tti_fast_marching_equals_jacobi(O):
    l, qm = hexagonal(**HEX)
    n = 24
    rng = np.random.default_rng(4)
    cost = 1 + 0.3 * rng.uniform(size=(n, n))
    P = O.make_problem('TTI', (n, n), [1. / n] * 2, metric=(l, qm, None), cost=cost, seeds_idx=[[5, 7]])
    u_fm = P.solve_fm()
    u_j, it = P.solve_jacobi(1e-14, 1e-14)
    np.testing.assert_allclose(u_j, u_fm, rtol=1e-10, atol=1e-12)


def test_tti3_without_quadratic_part_is_riemannian(O):
    n = 12
    h = np.full(3, 1. / n)
    l = np.array([1.5, 0.6])
    P_tti = O.make_problem('TTI', (n,) * 3, h, metric=(l, np.zeros((2, 2)), None), seeds_idx=[[3, 4, 5]])
    M = np.diag(1. / np.array([l[0], l[0], l[1]]))
    P_r = O.make_problem('Riemann', (n,) * 3, h, metric=M, seeds_idx=[[3, 4, 5]])
    np.testing.assert_allclose(P_tti.solve_fm(), P_r.solve_fm(), rtol=1e-10)


@pytest.mark.parametrize("d", [2, 3])
def test_tti_second_order_without_quadratic_part_is_riemannian(O, d):
    """The second-order correction applied to every sub-stencil: with q = 0
    the 7 sub-stencils are the Riemannian one, so the second-order TTI fixed
    point equals the second-order Riemannian one (itself pinned by its
    convergence rates), and differs from the first-order one."""
    n = 25 if d == 2 else 11
    h = np.array([1. / n] * d)
    if d == 2:
        l = np.array([2., 0.7])
        th = 0.4
        A = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]) @ np.diag([1., 1.5])
    else:
        l = np.array([2., 0.7])
        A = None                      # (TTI3: transversely isotropic about the last axis)
    seed = [[n // 3] * d]
    P_tti = O.make_problem('TTI', (n,) * d, h, metric=(l, np.zeros((2, 2)), A), seeds_idx=seed)
    if d == 2:
        M = A.T @ np.diag(1. / l) @ A
    else:
        M = np.diag(1. / np.array([l[0], l[0], l[1]]))
    P_r = O.make_problem('Riemann', (n,) * d, h, metric=M, seeds_idx=seed)
    with O.order(2):
        u_t, _ = P_tti.solve_jacobi(1e-13, 0.)
        u_r, _ = P_r.solve_jacobi(1e-13, 0.)
    np.testing.assert_allclose(u_t, u_r, rtol=1e-9)
    u1, _ = P_tti.solve_jacobi(1e-13, 0.)
    assert np.max(np.abs(u1 - u_t)) > 1e-4


@pytest.mark.parametrize("d", [2, 3])
def test_tti_geometry_of_field_inputs(O, d):
    """tti_geometry with linear, quadratic and transformation given as
    per-cell fields equals its constant form broadcast, cell by cell, and
    its last d*d components are (A diag(h))^{-T} (row-major)."""
    shape = (5, 4) if d == 2 else (3, 4, 2)
    h = np.array([0.1, 0.2, 0.05][:d])
    rng = np.random.default_rng(9)
    l = np.array([2., 0.7])
    q = np.array([[0.3, -0.1], [-0.1, 0.2]])
    A = np.linalg.qr(rng.normal(size=(d, d)))[0] * 1.3
    g_const = O.tti_geometry((l, q, A), h, shape)
    ext = (1,) * d
    field = lambda x: np.broadcast_to(x.reshape(x.shape + ext), x.shape + shape).copy()
    g_field = O.tti_geometry((field(l), field(q), field(A)), h, shape)
    assert g_field.shape == g_const.shape == (5 + d * d,) + shape
    assert np.array_equal(g_field, g_const)
    T = np.linalg.inv(A * h[None, :]).T
    assert np.allclose(g_const[5:].reshape((d, d) + shape)[(...,) + (0,) * d], T, rtol=1e-13)
