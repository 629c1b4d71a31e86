"""Pins the oracle's upwind update (HFM.h:142-207) with closed forms and an
independent bisection on the defining equation."""
import numpy as np
import pytest


def bisection_update(w, v, rhs):
    """Smallest u with sum_k w_k (u - v_k)_+^2 = rhs^2 (bisection)."""
    v = np.asarray(v, dtype=float)
    w = np.asarray(w, dtype=float)
    fin = np.isfinite(v) & (w > 0)
    if not np.any(fin):
        return np.inf
    lo = np.min(v[fin])
    hi = lo + rhs / np.sqrt(np.max(w[fin])) * 10 + 10 * rhs / np.sqrt(np.min(w[fin]))

    def f(u):
        return np.sum(w[fin] * np.maximum(0., u - v[fin]) ** 2) - rhs ** 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def test_single_neighbour(O):
    u = O.hfm_update([[4., 1.]], [[2., np.inf]], [3.])
    assert u[0] == pytest.approx(2. + 3. / 2.)


def test_far_point_stays_infinite(O):
    assert np.isinf(O.hfm_update([[1., 1.]], [[np.inf, np.inf]], [1.])[0])


def test_rouy_tourin_closed_form(O):
    """Isotropic 2D update with unit weights (teaching kernel
    Notebooks_NonDiv/TestData/IsoEik_GPU.h:52-64)."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        v0, v1 = rng.uniform(0, 2, 2)
        c = rng.uniform(0.1, 2)
        w0, w1 = min(v0, v1), max(v0, v1)
        ref = w0 + c
        if ref > w1:
            ref = (w0 + w1 + np.sqrt(max(0., 2 * c * c - (w0 - w1) ** 2))) / 2
        u = O.hfm_update([[1., 1.]], [[v0, v1]], [c])[0]
        assert u == pytest.approx(ref, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("nact", [2, 3, 6])
def test_matches_bisection(O, nact):
    rng = np.random.default_rng(nact)
    n = 300
    w = rng.uniform(0, 3, (n, nact)) * (rng.uniform(size=(n, nact)) > 0.2)
    w[:, 0] += 0.1
    v = rng.uniform(0, 1, (n, nact))
    v[rng.uniform(size=(n, nact)) < 0.15] = np.inf
    v[:, 0] = rng.uniform(0, 1, n)
    rhs = rng.uniform(0.05, 1, n)
    u, fw = O.hfm_update(w, v, rhs, flow=True)
    for i in range(n):
        ref = bisection_update(w[i], v[i], rhs[i])
        assert u[i] == pytest.approx(ref, rel=1e-10, abs=1e-12)
    # flow weights: fw_k = w_k (u - v_k)_+ and sum_k fw_k (u - v_k) = rhs^2
    diff = np.maximum(0., u[:, None] - np.where(np.isfinite(v), v, u[:, None]))
    np.testing.assert_allclose(fw, w * diff, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(np.sum(fw * diff, axis=1), rhs ** 2, rtol=1e-9)


def test_float_version_close(O):
    rng = np.random.default_rng(5)
    w = rng.uniform(0.1, 3, (500, 6))
    v = rng.uniform(0, 1, (500, 6))
    rhs = rng.uniform(0.05, 1, 500)
    np.testing.assert_allclose(O.hfm_update(w, v, rhs, use_float=True),
                               O.hfm_update(w, v, rhs), rtol=1e-5)
