"""Rander (drift) model: flow export and automatic differentiation, whose
flow is divided by its metric norm (_PostProcess.py:80-91, 185, 213-214).

* Exported flow: unit Rander norm at every cell off the seeds (the closed
  form of the normalisation: N(-h flow / N_h(-flow)) = 1 by homogeneity),
  and, for a constant metric, the direction of the straight geodesic from
  the seed.
* Forward AD (costVariation, at unit cost): the right-hand side carries the
  flow's metric norm N, which makes it the derivative of the solve with the
  Rander metric scaled by the cost (drift included): against central finite
  differences of the oracle's FP64 solve with metric (m (1 +- eps dc)^2,
  w (1 +- eps dc)) (eps = 1e-4).
* Reverse AD: the reference divides the adjoint by N where the forward
  multiplies by it, so costSensitivity * N^2 satisfies the adjoint identity
  <costSensitivity * N^2, dc> = <sensitivity, du> with du from the forward
  solve (unit cost).
"""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import Eikonal, Metrics, workloads
from gpu_helpers import log, oracle_problem

pytestmark = pytest.mark.gpu


def rander_problem(n=48, seed=21):
    wl = workloads.riemann2_aniso(n=n, seed=seed, sigma=4., nseeds=1, amax=3.)
    rng = np.random.default_rng(seed)
    drift = (0.3 * rng.uniform(-1, 1, (2,))[:, None, None] * np.ones((2,) + wl.dims)).astype(np.float32)
    wl.model = 'Rander2'
    wl.extra['drift'] = drift
    return wl


def rander_dict(wl, **extra):
    hfmIn = wl.dict_in(**extra)
    hfmIn['metric'] = Metrics.Rander(wl.metric, wl.extra['drift'])
    return hfmIn


def test_rander_flow_has_unit_norm():
    n = 81
    m = np.array([[1.5, 0.3], [0.3, 1.]], np.float32)
    w = np.array([0.35, -0.2], np.float32)
    hfmIn = Eikonal.dictIn({'model': 'Rander2', 'seed': [0., 0.], 'exportValues': True,
                            'exportGeodesicFlow': True, 'metric': Metrics.Rander(m, w), 'verbosity': 0})
    hfmIn.SetRect([[-1, 1], [-1, 1]], dims=[n, n])
    out = hfmIn.Run()
    flow = out['flow'].cpu().numpy().astype(np.float64)
    X = hfmIn.Grid().astype(np.float64)
    norm = np.sqrt(np.einsum('i...,ij,j...->...', flow, m.astype(np.float64), flow)) + \
        np.einsum('i,i...->...', w.astype(np.float64), flow)
    c = n // 2
    off = np.ones((n, n), bool)
    off[c, c] = False
    log(f"rander flow: Rander norm off the seed in [{norm[off].min():.6f}, {norm[off].max():.6f}]")
    assert np.allclose(norm[off], 1., atol=2e-5)
    # straight geodesics: the flow at x points along x - seed (away from it)
    r = X - X[:, c, c].reshape(2, 1, 1)
    far = np.sqrt((r ** 2).sum(0)) > 0.3
    cosang = (flow * r)[:, far].sum(0) / np.sqrt((flow[:, far] ** 2).sum(0) * (r[:, far] ** 2).sum(0))
    log(f"rander flow direction: median angle to x - seed {np.degrees(np.median(np.arccos(np.clip(cosang, -1, 1)))):.2f} deg")
    assert np.median(cosang) > np.cos(np.radians(5.))


def test_rander_forward_ad_matches_finite_differences(O):
    wl = rander_problem()
    rng = np.random.default_rng(5)
    dcost = (0.5 * rng.normal(size=wl.dims + (1,))).astype(np.float32)
    out = rander_dict(wl, exportValues=True, costVariation=dcost).Run()
    du = out['values'].coef.cpu().numpy()[..., 0].astype(np.float64)
    eps = 1e-4
    dc = dcost[..., 0].astype(np.float64)
    m0, w0 = wl.metric.astype(np.float64), wl.extra['drift'].astype(np.float64)

    def solve(scale):
        wl.metric = m0 * scale ** 2
        wl.extra['drift'] = w0 * scale
        P, _ = oracle_problem(O, wl, cast=False)
        u, _ = P.solve_jacobi(1e-13, 1e-13)
        return u

    fd = (solve(1 + eps * dc) - solve(1 - eps * dc)) / (2 * eps)
    err = np.max(np.abs(du - fd)) / np.max(np.abs(fd))
    log(f"rander forward AD vs FP64 finite differences of the scaled metric: rel err {err:.2e}")
    assert err <= 2e-2


def test_rander_reverse_ad_adjoint_identity():
    wl = rander_problem(seed=23)
    rng = np.random.default_rng(6)
    dcost = rng.normal(size=wl.dims + (1,)).astype(np.float32)
    sens = np.zeros(wl.dims + (1,), np.float32)
    for _ in range(12):
        sens[tuple(rng.integers(0, wl.dims[0], 2)) + (0,)] += 1.
    hfmIn = rander_dict(wl, exportValues=True, costVariation=dcost, sensitivity=sens, retself=True)
    out, iface = hfmIn.Run()
    du = out['values'].coef.cpu().numpy()[..., 0].astype(np.float64)
    lam = out['costSensitivity'].cpu().numpy()[..., 0].astype(np.float64)
    N = iface.flow_normalization.cpu().numpy().astype(np.float64)
    lhs = float((lam * N ** 2 * dcost[..., 0]).sum())
    rhs = float((sens[..., 0] * du).sum())
    log(f"rander adjoint identity: <lambda N^2, dc> = {lhs:.6f}, <s, du> = {rhs:.6f}")
    assert abs(lhs - rhs) <= 1e-3 * abs(rhs)
