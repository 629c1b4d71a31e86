"""Pins the oracle's second-order scheme (HFM.h:91-134): second-order
convergence to smooth exact solutions where the first-order scheme converges
at first order, and the first-order scheme where the criterion never holds."""
import numpy as np
import pytest


def disk_problem(O, n, M, R0=0.35, model='Riemann', order=1):
    """Exact solution sqrt(x^T M x) imposed on the cells with r < R0 (the seeds),
    solved outside: a smooth problem."""
    h = 2. / (n - 1)
    x = (np.arange(n) - n // 2) * h
    X = np.stack(np.meshgrid(x, x, indexing='ij'))
    ex = np.sqrt(np.einsum('i...,ij,j...->...', X, M, X))
    inner = ex < R0
    idx = np.argwhere(inner)
    kw = {'metric': M} if model == 'Riemann' else {}
    P = O.make_problem(model, (n, n), [h, h], seeds_idx=idx, seed_values=ex[inner], **kw)
    return P, ex, inner


@pytest.mark.parametrize("model,M", [("Isotropic", np.eye(2)),
                                     ("Riemann", np.array([[2., 0.6], [0.6, 1.]]))])
def test_order2_converges_at_second_order(O, model, M):
    """Mean error (the grid corners, where the further neighbours leave the
    grid, keep a first-order error that does not spread)."""
    errs = {1: [], 2: []}
    for k in (1, 2):
        for n in (41, 81, 161):
            P, ex, inner = disk_problem(O, n, M, model=model)
            with O.order(k):
                u, it = P.solve_jacobi(1e-13, 0.)
            errs[k].append(np.mean(np.abs(u - ex)))
    r1 = np.log2(errs[1][1] / errs[1][2])
    r2 = np.log2(errs[2][1] / errs[2][2])
    assert 0.6 < r1 < 1.4, errs[1]          # first order
    assert r2 > 1.6, errs[2]                # second order
    assert errs[2][2] < errs[1][2] / 5


def test_order2_fast_marching_refused(O):
    P, _, _ = disk_problem(O, 21, np.eye(2))
    with O.order(2):
        with pytest.raises(ValueError):
            P.solve_fm()


def test_order2_threshold_zero_is_first_order(O):
    """With threshold 0 the criterion |u-2v+v2| < 0 never holds."""
    P, ex, _ = disk_problem(O, 41, np.array([[2., 0.6], [0.6, 1.]]))
    u1, _ = P.solve_jacobi(1e-13, 0.)
    with O.order(2, threshold=0.):
        u2, _ = P.solve_jacobi(1e-13, 0.)
    np.testing.assert_array_equal(u1, u2)
