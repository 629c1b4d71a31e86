"""Pins the TTI source factorization of the oracle (TTI_.h:139-212 with
GoldenSearch.h and EuclideanFactor.h:34-50) and the host's TTI cost metric
and norm (_Interface.py:202-211, _SetArgs.py:46-78, Seismic/implicit_base.py).

The golden-section norm is checked against the support function of the dual
ball computed by brute force (a wrong ellipse parametrisation or a min/max
swap is off by far more than the search's own error), reduces exactly to the
elliptic norm when the quadratic part vanishes (where the shifts must equal
those of the closed-form Riemannian norm), and the factored scheme of a
constant medium must then be exact near the seed up to that search error."""
import numpy as np
import pytest

HEX = [dict(c11=15., c13=4., c33=10., c44=3.), dict(c11=1.5, c13=0.3, c33=1.0, c44=0.4)]


def hexagonal(c11, c13, c33, c44):
    """TTI.from_hexagonal (Metrics/Seismic/tti.py:141-146)."""
    linear = np.array([c11 + c44, c33 + c44])
    mixed = c13 ** 2 - c11 * c33 + 2. * c13 * c44
    quadratic = np.array([[-2. * c11 * c44, mixed], [mixed, -2. * c33 * c44]])
    return linear, quadratic


def q3_of(q):
    return np.array([q[0, 0], q[1, 0], q[1, 1]])


def transform(d, rng, scale=0.7):
    return np.linalg.qr(rng.normal(size=(d, d)))[0] * scale


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("hexp", HEX)
def test_golden_norm_matches_the_support_function(O, d, hexp):
    rng = np.random.default_rng(1 + d)
    l, q = hexagonal(**hexp)
    A = transform(d, rng)
    fm = np.concatenate([l, q3_of(q), A.reshape(-1)])
    x = rng.normal(size=(300, d)) * 5
    norm, grad = O.tti_factor(d, fm, x)
    exact = O.tti_primal_norm(l, q3_of(q), A, x)
    # 8 golden-section steps: about 2e-5 relative (measured 2.3e-5 at most)
    assert np.max(np.abs(norm / exact - 1)) < 5e-5
    # the gradient of an ellipse norm is exact for Euler's identity
    np.testing.assert_allclose(np.sum(grad * x, axis=1), norm, rtol=1e-12)
    # and close to the true gradient (central differences of the support function)
    eps = 1e-5
    gb = np.stack([(O.tti_primal_norm(l, q3_of(q), A, x + eps * np.eye(d)[c])
                    - O.tti_primal_norm(l, q3_of(q), A, x - eps * np.eye(d)[c])) / (2 * eps)
                   for c in range(d)], 1)
    assert np.max(np.abs(grad - gb)) < 1e-2 * np.max(np.abs(gb))


@pytest.mark.parametrize("d", [2, 3])
def test_golden_norm_without_quadratic_part_is_elliptic(O, d):
    """q = 0: every tangent ellipse is the dual ball itself, the norm is
    sqrt(<Ax, diag(1/l) Ax>) and the shifts are those of that Riemannian
    norm, T(x) - T(x -+ e) -+ <e, grad T(x)>, in closed form."""
    rng = np.random.default_rng(7)
    l = np.array([2., 0.7])
    A = transform(d, rng, 1.3)
    fm = np.concatenate([l, np.zeros(3), A.reshape(-1)])
    x = rng.normal(size=(50, d)) * 4
    e = rng.integers(-2, 3, size=(50, d))
    e[np.all(e == 0, axis=1), 0] = 1
    norm, grad, fact, fact2 = O.tti_factor(d, fm, x, e, order2=True)
    dl = np.array([l[0]] * (d - 1) + [l[1]])
    M = A.T @ np.diag(1. / dl) @ A
    T = lambda y: np.sqrt(np.einsum('ki,ij,kj->k', y, M, y))
    np.testing.assert_allclose(norm, T(x), rtol=1e-13)
    np.testing.assert_allclose(grad, (x @ M) / T(x)[:, None], rtol=1e-12, atol=1e-14)
    ge = np.sum((x @ M) * e, axis=1) / T(x)
    np.testing.assert_allclose(fact[:, 0], -ge + T(x) - T(x - e), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(fact[:, 1], ge + T(x) - T(x + e), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(fact2[:, 0], 2 * fact[:, 0] - (T(x) - 2 * T(x - e) + T(x - 2 * e)),
                               rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(fact2[:, 1], 2 * fact[:, 1] - (T(x) - 2 * T(x + e) + T(x + 2 * e)),
                               rtol=1e-9, atol=1e-12)
    # FP32 restatement close to FP64
    n32, _, f32 = O.tti_factor(d, fm, x, e, f32=True)
    np.testing.assert_allclose(n32, norm, rtol=1e-5)
    assert np.max(np.abs(f32 - fact)) < 1e-4 * np.max(norm)


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("hexp", HEX)
def test_host_tti_norm_matches_the_support_function(O, d, hexp):
    from adaptivegriddiscretizations_amd.Eikonal import _factor
    rng = np.random.default_rng(11 + d)
    l, q = hexagonal(**hexp)
    A = transform(d, rng)
    v = rng.normal(size=(400, d)) * 3
    v[0] = 0.
    got = _factor.tti_norm(l, q3_of(q), A, v)
    ref = O.tti_primal_norm(l, q3_of(q), A, v)
    assert got[0] == 0.
    np.testing.assert_allclose(got[1:], ref[1:], rtol=1e-9)


@pytest.mark.parametrize("d", [2, 3])
def test_host_tti_cost_metric_and_seed_spreading(O, d):
    """CostMetric('TTI') (l, q and A diag(h) cost, interpolated at order 1)
    against the oracle's, its factor metric, and the spread seed values."""
    from adaptivegriddiscretizations_amd.Eikonal import _factor
    rng = np.random.default_rng(5)
    shape = (9, 8) if d == 2 else (7, 6, 5)
    h = np.array([0.1, 0.12, 0.08][:d])
    l, q = hexagonal(**HEX[0])
    lf = l.reshape((2,) + (1,) * d) * (1 + 0.1 * rng.uniform(size=(2,) + shape))
    A = transform(d, rng, 1.1)
    cost = 1 + 0.3 * rng.uniform(size=shape)
    seed = np.array([3.4, 2.7, 1.2][:d])
    cm = _factor.CostMetric('TTI', h, cost, (lf, q, A))
    a, _ = cm.at(seed[None], shape, [False] * d)
    ref = O.cost_metric_at('TTI', seed[None], h, shape, metric=(lf, q, A), cost=cost)
    assert ref[0] == 'tti'
    for mine, theirs in zip(a, ref[1:]):
        np.testing.assert_allclose(mine, theirs, rtol=1e-13)
    np.testing.assert_allclose(cm.factor_metric(a), O.factor_metric(ref), rtol=1e-13)
    cells, vals = _factor.spread_seed(cm, seed, 2., shape, [False] * d, 0.5)
    cells_o, vals_o = O.spread_seed(seed, 2., shape, lambda p: O.cost_metric_at(
        'TTI', p, h, shape, metric=(lf, q, A), cost=cost), value=0.5)
    assert np.array_equal(cells, cells_o)
    np.testing.assert_allclose(vals, vals_o, rtol=1e-9)


@pytest.mark.parametrize("d,n", [(2, 41), (3, 17)])
def test_factored_tti_solve_is_exact_near_the_seed(O, d, n):
    """A constant tilted medium: inside the factoring ball the factored
    scheme reproduces the exact distance up to the golden-search error,
    where the plain scheme errs by the first-order consistency error near
    the source (measured: 2.6e-5 against 4.2e-3 in 2D, 6.3e-5 against
    1.6e-2 in 3D)."""
    l, q = hexagonal(**HEX[0])
    tilt = 0.5
    c, s_ = np.cos(tilt), np.sin(tilt)
    A = (np.array([[c, -s_], [s_, c]]) if d == 2
         else np.array([[1., 0., 0.], [0., c, -s_], [0., s_, c]]))
    h = np.full(d, 1. / n)
    s = n // 2
    seed = [s] * d
    P = O.make_problem('TTI', (n,) * d, h, metric=(l, q, A), seeds_idx=[seed])
    u0, _ = P.solve_jacobi(1e-12, 0.)
    fm = O.factor_metric(O.cost_metric_at('TTI', np.array([seed], float), h, (n,) * d, metric=(l, q, A)))
    R = 6
    with O.factor(R, seed, fm):
        u1, _ = P.solve_jacobi(1e-12, 0.)
    X = np.stack(np.meshgrid(*[np.arange(n) - s] * d, indexing='ij'), -1).reshape(-1, d).astype(float)
    exact = O.tti_primal_norm(l, q3_of(q), A * h[None, :], X).reshape((n,) * d)
    ball = (np.sqrt(np.sum(X ** 2, axis=1)) < R).reshape((n,) * d)
    e0 = np.max(np.abs(u0 - exact)[ball])
    e1 = np.max(np.abs(u1 - exact)[ball])
    assert e1 < 2e-4
    assert e1 < 0.02 * e0
