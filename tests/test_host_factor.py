"""Host-side seed spreading and factorization constants of the product
(Eikonal/_factor.py) against the oracle's restatement (oracle/oracle.py) on
random fields, without a GPU (numpy inputs take the host path)."""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd.Eikonal import _factor as F

SHAPE = (23, 17)
H = np.array([0.07, 0.11])


def fields(model, rng):
    n0, n1 = SHAPE
    cost = 1. + 0.3 * rng.uniform(size=SHAPE)
    kw_p, kw_o = {'cost': cost}, {'cost': cost}
    if model == 'Diagonal':
        c = 0.5 + rng.uniform(size=(2,) + SHAPE)
        kw_p['metric'] = kw_o['metric'] = c
    elif model in ('Riemann', 'Rander'):
        a = rng.normal(size=(2, 2) + SHAPE)
        m = np.einsum('ik...,jk...->ij...', a, a) + 0.5 * np.eye(2)[:, :, None, None]
        kw_p['metric'] = kw_o['metric'] = m
        if model == 'Rander':
            w = 0.1 * rng.normal(size=(2,) + SHAPE)
            kw_p['drift'] = kw_o['drift'] = w
    return kw_p, kw_o


@pytest.mark.parametrize("model", ['Isotropic', 'Diagonal', 'Riemann', 'Rander'])
@pytest.mark.parametrize("periodic", [(0, 0), (1, 0)])
def test_spread_seed_matches_oracle(O, model, periodic):
    rng = np.random.default_rng(len(model) + 7 * periodic[0])
    kw_p, kw_o = fields(model, rng)
    cm = F.CostMetric(model, H, kw_p['cost'], kw_p.get('metric'), kw_p.get('drift'))
    for seed in ([0.3, 8.6], [11.2, 7.9], [21.7, 0.4]):
        cells, vals = F.spread_seed(cm, np.array(seed), 2.5, SHAPE, periodic, 0.25)
        ocells, ovals = O.spread_seed(np.array(seed), 2.5, SHAPE,
                                      lambda p: O.cost_metric_at(model, p, H, SHAPE, periodic=list(periodic),
                                                                 **kw_o),
                                      value=0.25, periodic=list(periodic))
        got = {tuple(c): v for c, v in zip(cells.tolist(), vals)}
        ref = {tuple(c): v for c, v in zip(ocells.tolist(), ovals)}
        assert set(got) == set(ref)
        for c in ref:
            assert abs(got[c] - ref[c]) <= 1e-12 * max(1., abs(ref[c]))


@pytest.mark.parametrize("model", ['Isotropic', 'Diagonal', 'Riemann', 'Rander'])
def test_factor_metric_matches_oracle(O, model):
    rng = np.random.default_rng(3)
    kw_p, kw_o = fields(model, rng)
    cm = F.CostMetric(model, H, kw_p['cost'], kw_p.get('metric'), kw_p.get('drift'))
    seed = np.array([[9.37, 4.81]])
    a, _ = cm.at(seed, SHAPE, (0, 0))
    mine = np.asarray(cm.factor_metric(a))
    ref = np.asarray(O.factor_metric(O.cost_metric_at(model, seed, H, SHAPE, **kw_o)))
    np.testing.assert_allclose(mine, ref, rtol=1e-13)


def test_interpolation_reproduces_multilinear_fields():
    X = np.stack(np.meshgrid(np.arange(SHAPE[0]), np.arange(SHAPE[1]), indexing='ij'), -1).astype(float)
    f = 1. + 0.2 * X[..., 0] - 0.1 * X[..., 1] + 0.03 * X[..., 0] * X[..., 1]
    pts = np.array([[0., 0.], [3.25, 7.5], [21.9, 15.99], [22., 16.]])
    corners, weights = F.interpolation_stencil(pts, SHAPE, (0, 0))
    val = np.sum(weights * f[corners[..., 0], corners[..., 1]], axis=0)
    ex = 1. + 0.2 * pts[:, 0] - 0.1 * pts[:, 1] + 0.03 * pts[:, 0] * pts[:, 1]
    np.testing.assert_allclose(val, ex, rtol=1e-13)


def _reflect_spline1(values, pts, periodic):
    """The reference's order-1 spline evaluation with its 'reflect' boundary
    (Interpolation.py:113-147, spline_coefs returning the values themselves
    at order 1): base index floor(x), weights (1 - t, t) on (base, base + 1),
    indices wrapped modulo 2n then mirrored (i >= n -> 2n - 1 - i) on
    non-periodic axes, modulo n on periodic ones."""
    d = pts.shape[1]
    out = np.zeros(len(pts))
    b = np.floor(pts).astype(np.int64)
    t = pts - b
    for m in range(1 << d):
        hi = np.array([(m >> c) & 1 for c in range(d)])
        idx = b + hi
        w = np.prod(np.where(hi == 1, t, 1 - t), axis=1)
        for c in range(d):
            n = values.shape[c]
            if periodic[c]:
                idx[:, c] = np.mod(idx[:, c], n)
            else:
                i = np.mod(idx[:, c], 2 * n)
                idx[:, c] = np.where(i >= n, 2 * n - 1 - i, i)
        out += w * values[tuple(idx.T)]
    return out


@pytest.mark.parametrize("periodic", [(False, False), (True, False)])
def test_interpolation_matches_reference_reflect_mode(periodic):
    """The cost metric at an off-grid seed and its spread neighbours is
    interpolated at order 1: the product clamps positions onto the grid, the
    reference mirrors indices ('reflect').  For positions in [-1, n] on a
    non-periodic axis -- every seed and every spread neighbour lies in the
    box [-1/2, n - 1/2] -- both give the same value (index -1 mirrors to 0,
    index n to n - 1, each with its linear weight); beyond, they differ.
    Periodic axes wrap in both."""
    rng = np.random.default_rng(11)
    vals = rng.uniform(size=SHAPE)
    lo = np.array([-1. if not p else -7. for p in periodic])
    hi = np.array([n if not p else n + 7. for n, p in zip(SHAPE, periodic)])
    pts = lo + (hi - lo) * rng.uniform(size=(4000, 2))
    pts[:20] = np.array([[-1., 0.], [SHAPE[0], 3.], [-0.5, SHAPE[1] - 0.5], [0., SHAPE[1]]] * 5)
    corners, weights = F.interpolation_stencil(pts, SHAPE, periodic)
    ours = np.sum(weights * vals[tuple(np.moveaxis(corners, -1, 0))], axis=0)
    ref = _reflect_spline1(vals, pts, periodic)
    assert np.max(np.abs(ours - ref)) < 1e-12
    if not periodic[0]:
        far = np.array([[-1.6, 3.2], [SHAPE[0] + 0.7, 4.1]])
        c2, w2 = F.interpolation_stencil(far, SHAPE, periodic)
        ours2 = np.sum(w2 * vals[tuple(np.moveaxis(c2, -1, 0))], axis=0)
        assert np.max(np.abs(ours2 - _reflect_spline1(vals, far, periodic))) > 1e-3
