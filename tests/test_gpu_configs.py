"""The headline workloads at their full sizes on the GPU, against the oracle.

  C1  Riemann2 2000^2, 4 seeds, values and geodesic flow
  C2  Riemann3 200^3, 1 seed, 64 boundary receivers: values, and
      d(sum of receiver arrival times)/d(cost) through torch.autograd
      against the oracle's reverse solve and against central finite
      differences of FP64 solves (h = 1e-3) at 32 cells
  C3  256 batched isotropic problems (sides 64..1024) against their
      unbatched solves, four of them against the oracle

Bars: FP32 values within 1e-4 max u of the oracle's FP64 fast marching on the
same float32 inputs, identical finite masks; the FP32 - FP64 difference of
the GPU solves is logged per workload.  Each test states its own further
bars.  (The 512^2 isotropic workload runs in test_gpu_solver.py; the
learned warm start at 1024^2 in test_gpu_surrogate.py.)
"""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import FiniteDifferences as FD, workloads
from gpu_helpers import log, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

TOL32 = 1e-4


def _values(out):
    return out['values'].cpu().numpy().astype(np.float64)


def test_c1_riemann2_2000_values_and_flow(O):
    """Values against the oracle's fast marching; the flow against the
    oracle's flow pass from the same values.  The flow pass updates values
    (the reference's decreasing rule) while neighbouring tiles read them, so
    the GPU flow of a cell is that of the pre-pass values wherever the pass
    changed nothing within the cell's stencil reach.  The flow weights are
    differences u - v_k of values about h apart, so the FP32 rounding of u
    (one ulp is 2.4e-7 at u ~ 2.5), which differs between the device's
    square root and the oracle's correctly rounded one, is amplified by
    about 1/h: single-ulp differences show as ~1e-4 of max |flow| in
    isolated cells.  Bars (of max |flow|): away from changes, 99% of cells
    within 1e-6, 99.99% within 5e-4, all within 1e-3; near changes (where
    the pass's sub-tolerance value updates add to this), 99.9% within 3e-4.
    (Measured: 1.1e-7, 1.2e-4, 2.2e-4; 4.9e-5 and, in another run of the
    same build, 1.1e-4: the FP32 fixed point of the chaotic relaxation
    depends on the order in which tiles run, and so does the set of cells
    the flow pass still lowers.)"""
    from scipy.ndimage import binary_dilation
    from adaptivegriddiscretizations_amd.Eikonal import HFM_GPU
    wl = workloads.riemann2_aniso(n=2000)
    log("C1 riemann2 2000^2, 4 seeds: GPU FP32 solve + flow")
    it = HFM_GPU.Interface(wl.dict_in(exportValues=True, exportGeodesicFlow=True))
    it.SetKernelTraits()
    it.SetGeometry()
    it.SetArgs()
    it.Solve()
    u_pre = it.values.cpu().numpy().astype(np.float64)
    it.PostProcess()
    u_post = it.values.cpu().numpy().astype(np.float64)
    flow = it.hfmOut['flow'].cpu().numpy().astype(np.float64)
    st = it.stats['eikonal']
    log(f"  passes {st['niter_o']}, cell-updates {st['cell_updates']}")
    P, idx = oracle_problem(O, wl)
    assert len(idx) == 4
    ref = P.solve_fm()
    err = rel_err(u_post, ref)
    log(f"  FP32 vs oracle FP64 fast marching: rel err {err:.2e}")
    assert err <= TOL32

    u64 = _values(wl.dict_in(exportValues=True, traits={'Scalar': np.float64},
                             atol=1e-12, rtol=1e-12).Run())
    P64, _ = oracle_problem(O, wl, cast=False)
    err64 = rel_err(u64, P64.solve_fm())
    log(f"  FP64 GPU vs oracle FP64: rel err {err64:.2e}; FP32 vs FP64 GPU: {rel_err(u_post, u64):.2e}")
    assert err64 <= 1e-9

    fl = P.flow(u_pre, use_float=True)
    h = wl.h.reshape((2, 1, 1))
    ref_flow = -h * fl['vector']
    scale = np.max(np.abs(ref_flow))
    dev = np.max(np.abs(flow - ref_flow), axis=0) / scale
    changed = u_post != u_pre
    near = binary_dilation(changed, structure=np.ones((11, 11), bool))  # Selling offsets reach < 5 cells
    far = ~near
    qf = np.quantile(dev[far], [0.99, 0.9999])
    qn = np.quantile(dev[near], 0.999) if near.any() else 0.
    log(f"  flow pass changed {changed.sum()} cells; {near.sum()} cells within reach of a change")
    log(f"  flow away from changes: 99% {qf[0]:.2e}, 99.99% {qf[1]:.2e}, max {dev[far].max():.2e} "
        f"(of max |flow|); near changes: 99.9% {qn:.2e}, max {dev[near].max():.2e}")
    assert qf[0] <= 1e-6 and qf[1] <= 5e-4 and dev[far].max() <= 1e-3
    assert qn <= 3e-4
    # the pass's values from the pre-pass values, where no neighbour changed
    assert np.max(np.abs(fl['values'] - u_post)[far]) <= 1e-6 * np.max(np.abs(ref[np.isfinite(ref)]))


@pytest.fixture(scope="module")
def c2():
    wl = workloads.riemann3_aniso(n=200)
    return wl


def test_c2_riemann3_200_values(O, c2):
    wl = c2
    log("C2 riemann3 200^3: GPU FP32 and FP64 solves")
    out = wl.dict_in(exportValues=True).Run()
    u = _values(out)
    st = out['stats']['eikonal']
    log(f"  passes {st['niter_o']}, cell-updates {st['cell_updates']}")
    u64 = _values(wl.dict_in(exportValues=True, traits={'Scalar': np.float64},
                             atol=1e-12, rtol=1e-12).Run())
    log(f"  FP32 vs FP64 GPU: rel diff {rel_err(u, u64):.2e}")
    P, _ = oracle_problem(O, wl)
    log("  oracle fast marching (FP64, 8M cells)")
    ref = P.solve_fm()
    err = rel_err(u, ref)
    log(f"  FP32 vs oracle FP64 fast marching: rel err {err:.2e}")
    assert err <= TOL32


def _fp64_interface(wl):
    from adaptivegriddiscretizations_amd.Eikonal import HFM_GPU
    it = HFM_GPU.Interface(wl.dict_in(traits={'Scalar': np.float64}, atol=1e-12, rtol=1e-12))
    it.SetKernelTraits()
    it.SetGeometry()
    it.SetArgs()
    return it


def test_c2_riemann3_200_autograd(O, c2):
    """dJ/dcost for J = sum of the 64 receivers' arrival times, cost = 1:
    (i) FP32 autograd against the oracle's exact reverse solve on the same
    flow graph (1e-3 of max); (ii) FP64 autograd against central finite
    differences of FP64 solves (cost +- 1e-3 at one cell) at 24 cells drawn
    from the gradient's support and 8 uniform cells; (iii) FP32 against FP64
    autograd; (iv) the 1-homogeneity of J in the cost: sum cost * dJ/dcost = J."""
    import torch
    from adaptivegriddiscretizations_amd.Eikonal import autograd
    wl = c2
    rec = [tuple(int(x) for x in r) for r in wl.receivers]
    n = wl.dims[0]

    def grad_of(dtype):
        cost = torch.ones(wl.dims, device='cuda', dtype=dtype, requires_grad=True)
        extra = {'traits': {'Scalar': np.float64}, 'atol': 1e-12, 'rtol': 1e-12} \
            if dtype == torch.float64 else {}
        vals = autograd.eikonal_values(wl.dict_in(**extra), cost)
        J = sum(vals[r] for r in rec)
        J.backward()
        return (cost.grad.cpu().numpy().astype(np.float64), float(J.detach()),
                vals.detach().cpu().numpy().astype(np.float64))

    log("C2 autograd FP32 (64 receivers)")
    g32, J32, u32 = grad_of(torch.float32)
    log("C2 autograd FP64")
    g64, J64, _ = grad_of(torch.float64)
    gmax = np.max(np.abs(g64))
    log(f"  J32 {J32:.6f}, J64 {J64:.9f}, sum g64 {g64.sum():.9f}, FP32 vs FP64 grad "
        f"{np.max(np.abs(g32 - g64)) / gmax:.2e}")
    assert abs(g64.sum() - J64) <= 1e-6 * J64            # sum_c cost_c dJ/dcost_c = J (cost = 1)
    assert np.max(np.abs(g32 - g64)) <= 1e-3 * gmax

    log("  oracle flow and exact reverse solve on the FP32 values")
    P, idx = oracle_problem(O, wl)
    sens = np.zeros(wl.dims + (1,))
    for r in rec:
        sens[r + (0,)] += 1.
    lam = P.adjoint(u32, P.flow(u32, use_float=True), sens)[..., 0]
    for s in idx:
        lam[tuple(s)] = 0.
    err = np.max(np.abs(g32 - lam)) / np.max(np.abs(lam))
    log(f"  FP32 autograd vs oracle reverse solve: rel err {err:.2e}")
    assert err <= 1e-3

    # finite differences of FP64 solves, cost perturbed in the tiled rhs
    it = _fp64_interface(wl)
    shape_o = FD.round_up_ratio(np.array(it.shape), np.array(it.shape_i))
    rec_t = torch.tensor(FD.tiled_index(np.array(rec), it.shape_i, shape_o), device='cuda')
    h = it.solver_handle()
    from adaptivegriddiscretizations_amd import _lib

    def J_of():
        it.values_t.fill_(float('inf'))
        st = _lib.EikonalStats()
        _lib.check(h.solve(it._desc(), it._io(), st))
        return float(it.values_t[rec_t].sum())

    rng = np.random.default_rng(5)
    support = np.argwhere(np.abs(g64) > 1e-2 * gmax)
    cells = [tuple(int(x) for x in c) for c in support[rng.choice(len(support), 24, replace=False)]]
    cells += [tuple(int(x) for x in c) for c in rng.integers(0, n, (8, 3))]
    eps = 1e-3
    worst = 0.
    for c in cells:
        ti = int(FD.tiled_index(np.array(c), it.shape_i, shape_o))
        c0 = float(it.rhs_t[ti])
        it.rhs_t[ti] = c0 + eps
        jp = J_of()
        it.rhs_t[ti] = c0 - eps
        jm = J_of()
        it.rhs_t[ti] = c0
        fd = (jp - jm) / (2 * eps)
        e = abs(fd - g64[c]) / gmax
        worst = max(worst, e)
        log(f"  cell {c}: FD {fd:.6e}, autograd FP64 {g64[c]:.6e}, FP32 {g32[c]:.6e}, |diff|/max {e:.1e}")
    log(f"  worst |FD - autograd| / max|grad| = {worst:.2e}")
    assert worst <= 1e-3


def test_c3_batch256_matches_single_solves_and_oracle(O):
    """All 256 problems of the batched workload (sides 64..1024, 1-8 seeds),
    solved in one worklist, against their unbatched solves (same finite
    masks, values within 2e-5 max u: both stop at the same tolerance from
    different visit orders); four of them against the oracle (1e-4)."""
    import torch
    from adaptivegriddiscretizations_amd.Eikonal import batch
    wls = workloads.isotropic2_batch(ninst=256, device='cuda')
    hfms = [wl.dict_in(exportValues=True) for wl in wls]
    log(f"C3 batch of {len(hfms)}: {sum(w.ncells for w in wls) / 1e6:.1f} M cells")
    outs = batch.RunBatch(hfms)
    st = outs[0]['stats']['eikonal']
    log(f"  batched: passes {st['launches']}, gpu {st['gpu_ms']:.1f} ms")
    worst = 0.
    for i, (hfm, out) in enumerate(zip(hfms, outs)):
        ref = hfm.Run()['values'].cpu().numpy().astype(np.float64)
        u = out['values'].cpu().numpy().astype(np.float64)
        worst = max(worst, rel_err(u, ref))
    log(f"  worst rel diff batched vs single over 256 problems: {worst:.2e}")
    assert worst <= 2e-5
    for i in (0, 85, 170, 255):
        wl = wls[i]
        wl.cost = wl.cost.cpu().numpy()
        P, _ = oracle_problem(O, wl)
        err = rel_err(outs[i]['values'].cpu().numpy().astype(np.float64), P.solve_fm())
        log(f"  problem {i} {wl.dims}, {len(wl.seeds)} seeds: rel err vs oracle {err:.2e}")
        assert err <= TOL32
    torch.cuda.synchronize()
