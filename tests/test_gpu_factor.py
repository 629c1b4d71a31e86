"""Source factorization and seed spreading on the GPU (factoringRadius,
seedRadius; EuclideanFactor.h:8-30, FiniteDifferences.h:30-45,
_SetArgs.py:46-78): spread seeds and factorization constants against the
oracle's restatement, fixed points and flows against the oracle's factored
iteration, exactness for constant metrics, and seed sensitivities."""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import Eikonal, Metrics, workloads
from gpu_helpers import log, rel_err

pytestmark = pytest.mark.gpu
N = 64
SEED = (0.137, -0.211)


def fields(model, n=N):
    x = -1 + (np.arange(n) + 0.5) * (2. / n)
    X = np.stack(np.meshgrid(x, x, indexing='ij'))
    kw_gpu, kw_orc = {}, {}
    if model == 'Isotropic':
        cost = (1 + 0.3 * np.sin(2 * X[0]) * np.cos(3 * X[1])).astype(np.float32)
        kw_gpu['cost'] = cost
        kw_orc['cost'] = cost.astype(np.float64)
    elif model == 'Diagonal':
        costs = np.stack([1 + 0.3 * np.sin(2 * X[0] + X[1]), 1 + 0.2 * np.cos(X[0] - 2 * X[1])]).astype(np.float32)
        kw_gpu['metric'] = Metrics.Diagonal(costs)
        kw_orc['metric'] = costs.astype(np.float64)
    else:
        m = workloads.riemann2_aniso(n=n, seed=7, sigma=4., nseeds=1, amax=3.).metric
        kw_orc['metric'] = m.astype(np.float64)
        if model == 'Riemann':
            kw_gpu['metric'] = Metrics.Riemann(m)
        else:
            w = np.array([0.25, -0.15], dtype=np.float32)[:, None, None] * np.ones((2, n, n), np.float32)
            kw_gpu['metric'] = Metrics.Rander(m, w)
            kw_orc['drift'] = w.astype(np.float64)
    return kw_gpu, kw_orc


def gpu_run(model, n=N, **extra):
    kw_gpu, _ = fields(model, n)
    hfm = Eikonal.dictIn({'model': model + '2', 'seeds': [SEED], 'exportValues': True, 'verbosity': 0,
                          'retself': True, **kw_gpu, **extra})
    hfm.SetRect([[-1, 1], [-1, 1]], dims=[n, n])
    out, inst = hfm.Run()
    return out, inst


def oracle_setup(O, model, n=N, radius=2.):
    _, kw = fields(model, n)
    h32 = np.float32(2.) / np.float32(n)
    h = np.array([h32, h32], dtype=np.float64)
    seed32 = np.asarray(SEED, dtype=np.float32)
    cont = ((seed32 - (np.float32(-1.) + np.float32(0.5) * h32)) / h32).astype(np.float64)
    shape = (n, n)
    cm_at = lambda p: O.cost_metric_at(model, p, h, shape, **kw)
    cells, vals = O.spread_seed(cont, radius, shape, cm_at)
    P = O.make_problem(model, shape, h, seeds_idx=cells, seed_values=vals, **kw).cast(np.float32)
    m6 = np.asarray(O.factor_metric(cm_at(cont[None])), dtype=np.float32).astype(np.float64)
    return P, cont, cells, vals, m6


MODELS = ['Isotropic', 'Diagonal', 'Riemann', 'Rander']


@pytest.mark.parametrize("model", MODELS)
def test_spread_seeds_and_factor_constants_match_oracle(O, model):
    out, inst = gpu_run(model, factoringRadius=10.)
    P, cont, cells, vals, m6 = oracle_setup(O, model)
    got = {tuple(c): v for c, v in zip(inst.seedIndices.tolist(), inst.seedValues)}
    ref = {tuple(c): v for c, v in zip(cells.tolist(), vals)}
    assert set(got) == set(ref) and len(ref) >= 9
    for c in ref:
        assert abs(got[c] - ref[c]) <= 1e-5 * abs(ref[c]) + 1e-7, (c, got[c], ref[c])
    np.testing.assert_allclose(inst.factor['origin'], cont, rtol=1e-6)
    np.testing.assert_allclose(np.asarray(inst.factor['metric']), m6, rtol=1e-5)
    assert inst.factor['radius2'] == 100.


@pytest.mark.parametrize("model", MODELS)
def test_factored_values_match_oracle(O, model):
    out, inst = gpu_run(model, factoringRadius=10., atol=1e-7, rtol=0.)
    u = out['values'].cpu().numpy().astype(np.float64)
    P, cont, cells, vals, m6 = oracle_setup(O, model)
    with O.factor(10., cont, m6):
        ref, it = P.solve_jacobi(1e-12, 0.)
    ref0, _ = P.solve_jacobi(1e-12, 0.)
    err, gap = rel_err(u, ref), rel_err(ref0, ref)
    log(f"{model}2 {N}^2 factored: rel err vs oracle {err:.2e} (factored vs unfactored {gap:.2e}), "
        f"passes {out['stats']['eikonal']['launches']}")
    assert err <= 1e-5
    assert gap >= 20 * err      # the comparison resolves the factorization


@pytest.mark.parametrize("model", ['Isotropic', 'Riemann'])
def test_factored_order2_matches_oracle(O, model):
    # one sweep per visit: the schedule closest to the oracle's Jacobi
    # iteration.  The corrected scheme is not monotone, so the limit of the
    # decreasing iteration depends on the schedule (about 4e-3 here, the
    # factored shifts steepening the second differences near the seed).
    out, inst = gpu_run(model, factoringRadius=10., order=2, atol=1e-7, rtol=0., traits={'niter_i': 1})
    u = out['values'].cpu().numpy().astype(np.float64)
    P, cont, cells, vals, m6 = oracle_setup(O, model)
    with O.factor(10., cont, m6), O.order(2):
        ref, it = P.solve_jacobi(1e-12, 0.)
    err = rel_err(u, ref)
    log(f"{model}2 factored order 2: rel err vs oracle {err:.2e}")
    assert err <= 5e-3


def test_factored_flow_matches_oracle(O):
    out, inst = gpu_run('Riemann', factoringRadius=10., atol=1e-7, rtol=0., exportGeodesicFlow=True)
    u = out['values'].cpu().numpy().astype(np.float64)
    P, cont, cells, vals, m6 = oracle_setup(O, 'Riemann')
    with O.factor(10., cont, m6):
        fl = P.flow(u, use_float=True)
    h32 = float(np.float32(2.) / np.float32(N))
    flow = out['flow'].cpu().numpy()
    ref = -h32 * fl['vector']
    bad = np.mean(np.abs(flow - ref) > 1e-3 * np.max(np.abs(ref)))
    log(f"factored flow: fraction of cells off by >1e-3 max {bad:.2e}")
    assert bad <= 2e-3


@pytest.mark.parametrize("d", [2, 3])
def test_factored_constant_metric_is_exact(d):
    """Constant Riemannian metric, factorization over the whole grid: the
    solution is the distance |x - seed|_M up to float32 rounding (away from
    the border, where a Selling offset may lack its upwind neighbour)."""
    n = 96 if d == 2 else 40
    M = np.array([[2., 0.6], [0.6, 1.]]) if d == 2 else \
        np.array([[2., 0.5, 0.2], [0.5, 1.5, -0.3], [0.2, -0.3, 1.]])
    h = 2. / n
    seed = [0.05] * d
    errs = {}
    for radius in (0., 1e4):
        hfm = Eikonal.dictIn({'model': f'Riemann{d}', 'metric': Metrics.Riemann(M.astype(np.float32)),
                              'seeds': [seed], 'exportValues': True, 'verbosity': 0,
                              'factoringRadius': radius, 'seedRadius': 2., 'atol': 1e-7, 'rtol': 0.})
        hfm.SetRect([[-1, 1]] * d, dims=[n] * d)
        u = hfm.Run()['values'].cpu().numpy().astype(np.float64)
        x = -1 + (np.arange(n) + 0.5) * h
        X = np.stack(np.meshgrid(*[x] * d, indexing='ij'))
        diff = X - np.asarray(seed).reshape((d,) + (1,) * d)
        T = np.sqrt(np.einsum('i...,ij,j...->...', diff, M, diff))
        inner = tuple([slice(6, -6)] * d)
        errs[radius] = np.abs(u - T)[inner].max() / T.max()
    log(f"Riemann{d} constant metric: rel err unfactored {errs[0.]:.2e}, factored {errs[1e4]:.2e}")
    assert errs[1e4] <= 2e-5
    assert errs[0.] >= 1e-3


def test_spread_seed_sensitivity_sums_cells():
    """u shifts with the seed value: the seed-value sensitivity of any
    linear functional <s, u> is sum(s) (the spread cells' sensitivities
    summed), and the forward variation of u is 1 everywhere."""
    rng = np.random.default_rng(3)
    s = rng.uniform(0.5, 1.5, (N, N, 2)).astype(np.float32)
    out, inst = gpu_run('Isotropic', seedRadius=2., sensitivity=s, atol=1e-7, rtol=0.)
    ss = out['seedValueSensitivity'].cpu().numpy()
    assert ss.shape == (1, 2)
    np.testing.assert_allclose(ss[0], s.reshape(-1, 2).sum(0), rtol=1e-3)
    out, inst = gpu_run('Isotropic', seedRadius=2., seedValueVariation=np.ones((1, 1), np.float32),
                        atol=1e-7, rtol=0.)
    du = out['values'].coef
    du = du.cpu().numpy() if hasattr(du, 'cpu') else np.asarray(du)
    np.testing.assert_allclose(du, 1., rtol=1e-4)


def test_factoring_rejections():
    with pytest.raises(ValueError, match="single seed"):
        hfm = Eikonal.dictIn({'model': 'Isotropic2', 'seeds': [[0., 0.], [0.5, 0.5]],
                              'factoringRadius': 5., 'verbosity': 0})
        hfm.SetRect([[-1, 1], [-1, 1]], dims=[32, 32])
        hfm.Run()
