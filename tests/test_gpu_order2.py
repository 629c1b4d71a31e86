"""Second-order scheme on the GPU (order=2, HFM.h:91-134): fixed point and
flow against the oracle's second-order iteration, and second-order
convergence to a smooth exact solution.

The corrected scheme is not monotone, so the limit of the decreasing
iteration depends on the schedule: a cell relaxed several times against
neighbours that are still far from converged can undershoot, and the
decreasing update keeps the undershoot.  The oracle iterates globally
(Jacobi); the GPU comparisons therefore run one in-tile sweep per tile
visit (niter_i = 1), the schedule closest to it; with the default sweeps
the values stay within about 1e-2 of the oracle's."""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import Eikonal, Metrics
from gpu_helpers import log, rel_err

pytestmark = pytest.mark.gpu
M = np.array([[2., 0.6], [0.6, 1.]])


def disk(n, R0=0.35):
    h = 2. / (n - 1)
    x = (np.arange(n) - n // 2) * h
    X = np.stack(np.meshgrid(x, x, indexing='ij'))
    ex = np.sqrt(np.einsum('i...,ij,j...->...', X, M, X))
    inner = ex < R0
    return h, X, ex, inner


def gpu_solve(n, order, **extra):
    h, X, ex, inner = disk(n)
    seeds = np.stack([X[0][inner], X[1][inner]], -1).astype(np.float32)
    hfm = Eikonal.dictIn({'model': 'Riemann2', 'metric': Metrics.Riemann(M.astype(np.float32)),
                          'seeds': seeds, 'seedValues': ex[inner].astype(np.float32),
                          'order': order, 'exportValues': True, 'verbosity': 0, **extra})
    hfm.SetRect([[-1 - h / 2, 1 + h / 2]] * 2, dims=[n, n])
    out = hfm.Run()
    return out, h, ex, inner


def oracle(O, n, order):
    h, X, ex, inner = disk(n)
    hf = np.float32(2 + h) / np.float32(n)          # grid scale as the float32 dictIn computes it
    P = O.make_problem('Riemann', (n, n), [float(hf)] * 2, metric=M.astype(np.float32).astype(np.float64),
                       seeds_idx=np.argwhere(inner), seed_values=ex[inner].astype(np.float32))
    P = P.cast(np.float32)
    with O.order(order):
        u, _ = P.solve_jacobi(1e-12, 0.)
    return P, u


def test_order2_matches_oracle(O):
    n = 81
    out, h, ex, inner = gpu_solve(n, 2, atol=1e-7, rtol=0., traits={'niter_i': 1})
    u = out['values'].cpu().numpy().astype(np.float64)
    P, ref = oracle(O, n, 2)
    err = rel_err(u, ref)
    e_gpu, e_orc = np.mean(np.abs(u - ex)), np.mean(np.abs(ref - ex))
    log(f"order 2 {n}^2: rel err vs oracle {err:.2e}, mean error vs exact {e_gpu:.2e} "
        f"(oracle {e_orc:.2e}), passes {out['stats']['eikonal']['launches']}")
    assert err <= 1e-3
    assert e_gpu <= 1.5 * e_orc
    out8, *_ = gpu_solve(n, 2, atol=1e-7, rtol=0.)        # default sweeps
    err8 = rel_err(out8['values'].cpu().numpy().astype(np.float64), ref)
    log(f"order 2 {n}^2 default niter_i: rel err vs oracle {err8:.2e}")
    assert err8 <= 2e-2


def test_order2_flow_matches_oracle(O):
    n = 61
    out, h, ex, inner = gpu_solve(n, 2, atol=1e-7, rtol=0., exportGeodesicFlow=True)
    u = out['values'].cpu().numpy().astype(np.float64)
    P, _ = oracle(O, n, 2)
    with O.order(2):
        fl = P.flow(u, use_float=True)
    hf = np.float32(2 + h) / np.float32(n)
    flow = out['flow'].cpu().numpy()
    ref = -float(hf) * fl['vector']
    bad = np.mean(np.abs(flow - ref) > 1e-3 * np.max(np.abs(ref)))
    log(f"order 2 flow: fraction of cells off by >1e-3 max {bad:.2e}")
    assert bad <= 2e-3


def test_order2_converges_at_second_order():
    errs = {}
    for order in (1, 2):
        errs[order] = []
        for n in (81, 161):
            out, h, ex, inner = gpu_solve(n, order, atol=1e-8, rtol=0., traits={'niter_i': 1})
            u = out['values'].cpu().numpy().astype(np.float64)
            errs[order].append(np.mean(np.abs(u - ex)))
    r1 = np.log2(errs[1][0] / errs[1][1])
    r2 = np.log2(errs[2][0] / errs[2][1])
    log(f"mean errors order 1 {errs[1]} (rate {r1:.2f}), order 2 {errs[2]} (rate {r2:.2f})")
    assert r2 > 1.6 and r1 < 1.4


@pytest.mark.parametrize("case", ["riemann3", "rander2"])
def test_order2_is_a_fixed_point_of_the_oracle_update(O, case):
    """Second order on the 3D Riemannian and the 2D Rander models: the
    device values are a fixed point of the oracle's FP32 second-order update
    on the same FP32 inputs (its iteration started there stays there).  The
    limit reached from +inf depends on the schedule (tile Gauss-Seidel vs
    global Jacobi, one sweep per visit: 4.5e-3 apart on riemann3 24^3,
    9.8e-3 on rander2 48^2, while the oracle's own FP32 and FP64 limits
    differ by up to 1e-3), so that comparison is bounded by 2e-2; the device
    values are not the first-order solution."""
    from adaptivegriddiscretizations_amd import workloads
    from gpu_helpers import oracle_problem
    if case == "riemann3":
        wl = workloads.riemann3_aniso(n=24, seed=4, sigma=3., nreceivers=4)
    else:
        from test_gpu_rander import rander_problem
        wl = rander_problem(n=48, seed=25)
        wl.cost = None
    hfm = wl.dict_in(exportValues=True, order=2, atol=1e-7, rtol=0., traits={'niter_i': 1})
    if case == "rander2":
        hfm['metric'] = Metrics.Rander(wl.metric, wl.extra['drift'])
    out = hfm.Run()
    u = out['values'].cpu().numpy().astype(np.float64)
    P, _ = oracle_problem(O, wl)
    with O.order(2):
        fix, _ = P.solve_jacobi(1e-7, 0., u0=u, use_float=True)
        ref, _ = P.solve_jacobi(1e-7, 0., use_float=True)
        ref64, _ = P.solve_jacobi(1e-12, 0.)
    ref1, _ = P.solve_jacobi(1e-12, 0.)
    err_fix, err = rel_err(u, fix), rel_err(u, ref)
    log(f"{case} order 2: oracle iteration from the device values moves them by {err_fix:.2e}; "
        f"rel err vs oracle FP32 from +inf {err:.2e} (FP64 {rel_err(u, ref64):.2e}, FP32/FP64 spread "
        f"{rel_err(ref, ref64):.2e}); first order differs by {rel_err(ref1, ref64):.2e}")
    assert err_fix <= 1e-5
    assert err <= 2e-2
    assert rel_err(u, ref1) > 1e-2
