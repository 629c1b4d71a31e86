"""Parity of the flow export and of the reverse-mode sensitivities with the
oracle, plus the reference's stored known answer for seedValueSensitivity."""
import json
import os

import numpy as np
import pytest

from adaptivegriddiscretizations_amd import workloads, Eikonal
from gpu_helpers import log, oracle_problem

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def frac_bad(a, b, rtol):
    scale = np.max(np.abs(b))
    return np.mean(np.abs(a - b) > rtol * scale)


@pytest.mark.parametrize("which", ["riemann2", "riemann3", "isotropic2"])
def test_flow_matches_oracle(O, which):
    if which == "riemann2":
        wl = workloads.riemann2_aniso(n=96, seed=31, sigma=4., amax=6.)
    elif which == "riemann3":
        wl = workloads.riemann3_aniso(n=20, seed=32, sigma=3.)
    else:
        wl = workloads.isotropic2_bumps(n=96, seed=33)
    hfmIn = wl.dict_in(exportValues=True, exportGeodesicFlow=True)
    out = hfmIn.Run()
    u = out['values'].cpu().numpy().astype(np.float64)
    flow = out['flow'].cpu().numpy().astype(np.float64)
    P, _ = oracle_problem(O, wl)
    ref = P.flow(u, use_float=True)
    h = wl.h.reshape((wl.ndim,) + (1,) * wl.ndim)
    ref_flow = -h * ref['vector']
    bad = frac_bad(flow, ref_flow, 1e-3)
    log(f"flow {which}: fraction of cells off by >1e-3 max: {bad:.2e}")
    assert bad <= 2e-3
    # values after the flow pass agree with the oracle's pass
    assert np.max(np.abs(u - ref['values'])) <= 1e-5 * np.max(np.abs(u[np.isfinite(u)]))


def test_seed_value_sensitivity_kat():
    k = json.load(open(os.path.join(HERE, "golden", "ad_kat.json")))["seed_value_sensitivity"]
    hfmIn = Eikonal.dictIn({'model': k['model'], 'seeds': k['seeds'], 'exportValues': 1,
                            'traits': {'shape_i': tuple(k['shape_i']), 'niter_i': k['niter_i']},
                            'cost': k['cost'], 'verbosity': 0})
    hfmIn.SetRect(k['sides'], dimx=k['dimx'], sampleBoundary=k['sampleBoundary'])
    sens = np.zeros(hfmIn.shape + (1,), dtype=np.float32)
    sens[tuple(k['sensitivity_cell'])] = 1.
    hfmIn['sensitivity'] = sens
    out = hfmIn.Run()
    svs = out['seedValueSensitivity'].cpu().numpy()
    log(f"seedValueSensitivity {svs.tolist()}")
    assert np.max(np.abs(svs - np.array(k['expected']))) < k['tol']


@pytest.mark.parametrize("which", ["riemann2", "riemann3"])
def test_adjoint_matches_oracle(O, which):
    if which == "riemann2":
        wl = workloads.riemann2_aniso(n=80, seed=41, sigma=4., amax=6., nseeds=2)
    else:
        wl = workloads.riemann3_aniso(n=20, seed=42, sigma=3., nreceivers=16)
    rng = np.random.default_rng(0)
    sens = np.zeros(wl.dims + (2,), dtype=np.float32)
    if wl.receivers is not None:
        for r in wl.receivers:
            sens[tuple(r) + (0,)] += 1.
    else:
        cells = rng.integers(0, wl.dims[0], (16, wl.ndim))
        for c in cells:
            sens[tuple(c) + (0,)] += 1.
    sens[..., 1] = rng.uniform(size=wl.dims)
    hfmIn = wl.dict_in(exportValues=True, sensitivity=sens)
    out = hfmIn.Run()
    u = out['values'].cpu().numpy().astype(np.float64)
    lam = out['costSensitivity'].cpu().numpy().astype(np.float64)
    P, idx = oracle_problem(O, wl)
    fl = P.flow(u, use_float=True)
    lam_ref = P.adjoint(u, fl, sens.astype(np.float64))
    for s in idx:
        lam_ref[tuple(s)] = 0.
    for r in range(2):
        err = np.max(np.abs(lam[..., r] - lam_ref[..., r])) / np.max(np.abs(lam_ref[..., r]))
        log(f"adjoint {which} rhs {r}: rel err {err:.2e}, stats {out['stats']['reverseAD']}")
        assert err <= 1e-3


def _forward_rhs(cost, dcost, seeds_idx, dseed):
    r = (cost[..., None] * dcost).astype(np.float64)
    for s, d in zip(seeds_idx, dseed):
        r[tuple(s)] = d
    return r


@pytest.mark.parametrize("which", ["riemann2", "riemann3"])
def test_forward_ad_matches_oracle(O, which):
    """values.coef of a costVariation + seedValueVariation solve against the
    oracle's exact forward solve on the same flow graph."""
    from adaptivegriddiscretizations_amd import AutomaticDifferentiation as ad
    if which == "riemann2":
        wl = workloads.riemann2_aniso(n=80, seed=43, sigma=4., amax=6., nseeds=2)
    else:
        wl = workloads.riemann3_aniso(n=20, seed=44, sigma=3.)
    rng = np.random.default_rng(1)
    cost = (1 + 0.3 * rng.uniform(size=wl.dims)).astype(np.float32)
    wl.cost = cost
    nseeds = len(wl.seeds)
    dcost = rng.normal(size=wl.dims + (2,)).astype(np.float32)
    dseed = rng.normal(size=(2, nseeds)).astype(np.float32)  # (nfwd, nseeds), as the key expects
    hfmIn = wl.dict_in(exportValues=True, costVariation=dcost, seedValueVariation=dseed)
    out = hfmIn.Run()
    assert ad.is_ad(out['values'])
    u = out['values'].value.cpu().numpy().astype(np.float64)
    du = out['values'].coef.cpu().numpy().astype(np.float64)
    P, idx = oracle_problem(O, wl)
    fl = P.flow(u, use_float=True)
    r = _forward_rhs(cost.astype(np.float64), dcost.astype(np.float64), idx, dseed.T.astype(np.float64))
    du_ref = P.forward_ad(u, fl, r)
    for q in range(2):
        err = np.max(np.abs(du[..., q] - du_ref[..., q])) / np.max(np.abs(du_ref[..., q]))
        log(f"forward AD {which} channel {q}: rel err {err:.2e}, stats {out['stats']['forwardAD']}")
        assert err <= 1e-3


def test_forward_ad_homogeneity():
    """Scaling cost and seed values together scales u: the variation along
    (cost, seedValues) is u itself."""
    wl = workloads.riemann2_aniso(n=96, seed=45, sigma=4., amax=6., nseeds=2)
    rng = np.random.default_rng(2)
    cost = (1 + 0.5 * rng.uniform(size=wl.dims)).astype(np.float32)
    wl.cost = cost
    sv = np.array([0.1, 0.3], dtype=np.float32)
    hfmIn = wl.dict_in(exportValues=True, seedValues=sv, costVariation=cost[..., None],
                       seedValueVariation=sv[None, :])
    out = hfmIn.Run()
    u = out['values'].value.cpu().numpy().astype(np.float64)
    du = out['values'].coef.cpu().numpy()[..., 0].astype(np.float64)
    err = np.max(np.abs(du - u)) / np.max(u)
    log(f"forward AD homogeneity: max|du-u|/max u = {err:.2e}")
    assert err <= 1e-3


def test_autograd_cost_gradient(O):
    """torch.autograd through the solve: d(sum of receiver values)/d cost
    against the oracle's reverse solve (cost * lambda) on the same flow
    graph, and against central finite differences of the FP64 oracle solve
    (h = 1e-3) at 32 random cells."""
    import torch
    from adaptivegriddiscretizations_amd.Eikonal import autograd
    wl = workloads.riemann2_aniso(n=64, seed=46, sigma=4., amax=6., nseeds=1)
    rng = np.random.default_rng(3)
    cost_np = (1 + 0.3 * rng.uniform(size=wl.dims)).astype(np.float32)
    receivers = [tuple(int(x) for x in rng.integers(0, 64, 2)) for _ in range(8)]
    hfmIn = wl.dict_in()
    cost = torch.tensor(cost_np, device='cuda', requires_grad=True)
    values = autograd.eikonal_values(hfmIn, cost)
    J = sum(values[r] for r in receivers)
    J.backward()
    grad = cost.grad.cpu().numpy().astype(np.float64)
    u = values.detach().cpu().numpy().astype(np.float64)

    wl.cost = cost_np
    P, idx = oracle_problem(O, wl)
    sens = np.zeros(wl.dims + (1,))
    for r in receivers:
        sens[r + (0,)] += 1.
    lam = P.adjoint(u, P.flow(u, use_float=True), sens)[..., 0]
    for s in idx:
        lam[tuple(s)] = 0.
    g_ref = cost_np.astype(np.float64) * lam
    err = np.max(np.abs(grad - g_ref)) / np.max(np.abs(g_ref))
    log(f"autograd vs oracle reverse: rel err {err:.2e}")
    assert err <= 1e-3

    def J_fp64(c):
        wl.cost = c
        P2, _ = oracle_problem(O, wl, cast=False)
        uu = P2.solve_fm()
        return sum(uu[r] for r in receivers)

    seed_cells = {tuple(s) for s in idx}
    cells = [tuple(int(x) for x in c) for c in rng.integers(0, 64, (32, 2))]
    eps = 1e-3
    ok = 0
    for c in cells:
        if c in seed_cells:
            ok += 1
            continue
        cp, cm = cost_np.astype(np.float64), cost_np.astype(np.float64)
        cp[c] += eps
        cm[c] -= eps
        fd = (J_fp64(cp) - J_fp64(cm)) / (2 * eps)
        ok += abs(fd - grad[c]) <= 2e-2 * max(abs(fd), 1e-2 * np.max(np.abs(g_ref)))
    log(f"autograd vs FP64 finite differences: {ok}/32 cells within 2e-2")
    assert ok >= 30


def test_forward_ad_through_speed():
    """A denseAD 'speed' gives the variation of the solve with cost = 1/speed:
    the same coefficients as costVariation = -dspeed / speed^2."""
    from adaptivegriddiscretizations_amd import AutomaticDifferentiation as ad
    wl = workloads.riemann2_aniso(n=64, seed=47, sigma=4., amax=4., nseeds=1)
    rng = np.random.default_rng(4)
    speed = (1 + 0.5 * rng.uniform(size=wl.dims)).astype(np.float32)
    dspeed = rng.normal(size=wl.dims + (2,)).astype(np.float32)
    h1 = wl.dict_in(exportValues=True)
    h1['speed'] = ad.Dense.denseAD(speed, dspeed)
    du1 = h1.Run()['values'].coef.cpu().numpy().astype(np.float64)
    cost = (1. / speed).astype(np.float32)
    h2 = wl.dict_in(exportValues=True, cost=cost,
                    costVariation=(-dspeed / (speed ** 2)[..., None]).astype(np.float32))
    du2 = h2.Run()['values'].coef.cpu().numpy().astype(np.float64)
    err = np.max(np.abs(du1 - du2)) / np.max(np.abs(du2))
    log(f"forward AD through speed vs costVariation: rel diff {err:.2e}")
    assert err <= 1e-4
