"""Parity of the HIP eikonal solver with the CPU oracle (FP64 fast marching
of the same scheme on the same float32 inputs).

Bar: max|u_gpu - u_oracle| <= 1e-4 max u for FP32 at the default stopping
tolerance (the headline tolerance), and <= 1e-9 max u for the FP64 path;
finite/infinite masks identical.
"""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import Eikonal, Metrics, workloads
from gpu_helpers import log, oracle_problem, rel_err, run_gpu

pytestmark = pytest.mark.gpu

TOL32 = 1e-4


@pytest.mark.parametrize("solver", ["FIM", "AGSI", "global_iteration"])
def test_isotropic2_bumps(O, solver):
    wl = workloads.isotropic2_bumps(n=128, seed=3)
    log(f"isotropic2_bumps 128^2 solver={solver}")
    u, out = run_gpu(wl, solver=solver)
    P, _ = oracle_problem(O, wl)
    ref = P.solve_fm()
    err = rel_err(u, ref)
    log(f"  niter_o={out['stats']['eikonal']['niter_o']} rel err {err:.2e}")
    assert err <= TOL32


def test_isotropic2_512_seed_on_tie(O):
    """The 512^2 isotropic workload: seed at the origin lands on the rounding
    tie 255.5 -> 256 (half to even)."""
    wl = workloads.isotropic2_bumps(n=512, seed=0)
    u, out = run_gpu(wl, atol=1e-8, rtol=0.)
    P, idx = oracle_problem(O, wl)
    assert idx.tolist() == [[256, 256]]
    assert u[256, 256] == 0.
    ref = P.solve_fm()
    err = rel_err(u, ref)
    log(f"isotropic2 512^2 niter_o={out['stats']['eikonal']['niter_o']} rel err {err:.2e}")
    assert err <= TOL32


@pytest.mark.parametrize("n,solver,pre", [(128, "FIM", True), (150, "AGSI", True), (128, "FIM", False)])
def test_riemann2_aniso(O, n, solver, pre):
    wl = workloads.riemann2_aniso(n=n, seed=5, sigma=4., amax=10.)
    log(f"riemann2_aniso {n}^2 solver={solver} precompute_scheme={pre}")
    u, out = run_gpu(wl, solver=solver, precompute_scheme=pre)
    P, _ = oracle_problem(O, wl)
    ref = P.solve_fm()
    err = rel_err(u, ref)
    log(f"  niter_o={out['stats']['eikonal']['niter_o']} rel err {err:.2e}")
    assert err <= TOL32


def test_riemann2_float64_tight(O):
    wl = workloads.riemann2_aniso(n=96, seed=6, sigma=4., amax=6.)
    u, out = run_gpu(wl, traits={'Scalar': np.float64, 'shape_i': (8, 8), 'niter_i': 8},
                     atol=1e-13, rtol=1e-13)
    P, _ = oracle_problem(O, wl, cast=False)
    ref = P.solve_fm()
    err = rel_err(u, ref)
    log(f"riemann2 fp64 96^2 rel err {err:.2e}")
    assert err <= 1e-9


@pytest.mark.parametrize("pre", [True, False])
@pytest.mark.parametrize("dims", [(24, 24, 24), (30, 17, 21)])
def test_riemann3_aniso(O, dims, pre):
    wl = workloads.riemann3_aniso(n=dims[0], seed=4, sigma=3.)
    if dims != (dims[0],) * 3:
        wl.metric = np.ascontiguousarray(wl.metric[:, :, :dims[0], :dims[1], :dims[2]])
        wl.dims = dims
        wl.sides = [[-1, -1 + 2. * d / dims[0]] for d in dims]
    log(f"riemann3_aniso {dims} precompute_scheme={pre}")
    u, out = run_gpu(wl, precompute_scheme=pre)
    P, _ = oracle_problem(O, wl)
    ref = P.solve_fm()
    err = rel_err(u, ref)
    log(f"  niter_o={out['stats']['eikonal']['niter_o']} rel err {err:.2e}")
    assert err <= TOL32


def test_multiple_seeds_and_values(O):
    wl = workloads.riemann2_aniso(n=80, seed=9, sigma=3., nseeds=3, amax=4.)
    vals = np.array([0., 0.3, 0.1], dtype=np.float32)
    u, _ = run_gpu(wl, seedValues=vals)
    P, _ = oracle_problem(O, wl, seed_values=vals.astype(np.float64))
    assert rel_err(u, P.solve_fm()) <= TOL32


def test_walls(O):
    wl = workloads.isotropic2_bumps(n=96, seed=2)
    walls = np.zeros(wl.dims, dtype=bool)
    walls[40:44, 10:80] = True
    walls[60:90, 50:53] = True
    u, _ = run_gpu(wl, walls=walls)
    P, _ = oracle_problem(O, wl, walls=walls)
    ref = P.solve_fm()
    assert np.all(np.isinf(u[walls]))
    assert rel_err(u, ref) <= TOL32


@pytest.mark.parametrize("pre", [True, False])
def test_riemann2_walls_visibility(O, pre):
    wl = workloads.riemann2_aniso(n=96, seed=12, sigma=4., nseeds=1, amax=8.)
    walls = np.zeros(wl.dims, dtype=bool)
    walls[30:33, 5:70] = True
    u, _ = run_gpu(wl, walls=walls, precompute_scheme=pre)
    P, _ = oracle_problem(O, wl, walls=walls)
    assert rel_err(u, P.solve_fm()) <= TOL32


def test_rander2_matches_fp64_jacobi(O):
    wl = workloads.riemann2_aniso(n=64, seed=13, sigma=4., nseeds=1, amax=3.)
    rng = np.random.default_rng(1)
    drift = (0.3 * rng.uniform(-1, 1, (2,))[:, None, None] * np.ones((2,) + wl.dims)).astype(np.float32)
    wl.model = 'Rander2'
    wl.extra['drift'] = drift
    hfmIn = wl.dict_in(exportValues=True)
    hfmIn['metric'] = Metrics.Rander(wl.metric, drift)
    u = hfmIn.Run()['values'].cpu().numpy().astype(np.float64)
    P, _ = oracle_problem(O, wl)
    ref, it = P.solve_jacobi(1e-13, 1e-13)
    log(f"rander2 64^2 jacobi iterations {it}")
    assert rel_err(u, ref) <= TOL32


def test_nonconvergence_raises():
    wl = workloads.isotropic2_bumps(n=64, seed=1)
    with pytest.raises(ValueError, match="convergence"):
        run_gpu(wl, nitermax_o=1, check_interval=1)


def test_warm_start_supersolution_same_fixed_point(O):
    wl = workloads.riemann2_aniso(n=96, seed=21, sigma=4., amax=5.)
    u_cold, out_cold = run_gpu(wl)
    u0 = (1.05 * u_cold).astype(np.float32)
    u_warm, out_warm = run_gpu(wl, values=u0, warmStartAllTiles=True)
    P, _ = oracle_problem(O, wl)
    ref = P.solve_fm()
    assert rel_err(u_warm, ref) <= TOL32
    assert rel_err(u_cold, ref) <= TOL32


@pytest.mark.parametrize("graph", [False, True])
def test_graph_replay_matches_eager_launches(O, graph):
    """Pass chunks replayed from a captured hipGraph and eager launches reach
    the same fixed point (several replays per solve, two solves in a row)."""
    wl = workloads.riemann2_aniso(n=100, seed=17, sigma=4., amax=6.)
    P, _ = oracle_problem(O, wl)
    ref = P.solve_fm()
    for _ in range(2):
        u, out = run_gpu(wl, hipGraph=graph, check_interval=3)
        assert out['stats']['eikonal']['launches'] >= 6
        assert rel_err(u, ref) <= TOL32


@pytest.mark.parametrize("d", [2, 3])
def test_diagonal_matches_oracle(O, d):
    """Diagonal metrics (per-axis costs): Isotropic_.h with geometry."""
    from adaptivegriddiscretizations_amd import Metrics
    n = 80 if d == 2 else 20
    rng = np.random.default_rng(21 + d)
    costs = (1 + 0.5 * rng.uniform(size=(d,) + (n,) * d)).astype(np.float32)
    hfm = Eikonal.dictIn({'model': f'Diagonal{d}', 'metric': Metrics.Diagonal(costs),
                          'seeds': [[0.1] * d], 'exportValues': True, 'verbosity': 0})
    hfm.SetRect([[-1, 1]] * d, dims=[n] * d)
    u = hfm.Run()['values'].cpu().numpy().astype(np.float64)
    h = np.full(d, np.float32(2. / n)).astype(np.float64)
    idx = O.seed_indices([[0.1] * d], np.full(d, -1.), h)
    P = O.make_problem('Diagonal', (n,) * d, h, metric=costs.astype(np.float64), seeds_idx=idx)
    err = rel_err(u, P.cast(np.float32).solve_fm())
    log(f"diagonal{d}: rel err {err:.2e}")
    assert err <= TOL32


@pytest.mark.parametrize("key", ["multiprecision", "values_float64"])
def test_multiprecision_reaches_fp64_accuracy(O, key):
    """multiprecision / values_float64 run in FP64 arithmetic: with a tight
    tolerance the values match the exact fixed point far below FP32
    resolution (values_float64 also exports float64)."""
    wl = workloads.riemann2_aniso(n=96, seed=6, sigma=4., amax=6.)
    u, out = run_gpu(wl, atol=1e-11, **{key: True})
    assert out['values'].dtype == (torch_dtype('float64') if key == 'values_float64' else torch_dtype('float32'))
    P, _ = oracle_problem(O, wl, cast=False)
    err = rel_err(u, P.solve_fm())
    log(f"{key}: rel err {err:.2e}")
    assert err <= (1e-9 if key == 'values_float64' else 1e-7)


def torch_dtype(name):
    import torch
    return getattr(torch, name)


def test_grid_beyond_4gib_fields(O):
    """Maximum sizes: a 1040^3 isotropic grid (1.125e9 cells, 4.5 GB per FP32
    field) exceeds 32-bit byte offsets, so every pass runs the general
    variant with 64-bit addressing.  The discrete solution of a point source
    with unit cost is exact along the axes through the seed, symmetric under
    axis permutations, and its first-order error is below the one the
    oracle's fast marching shows on a 65^3 grid."""
    import torch
    n = 1040
    h = 2. / n
    c = n // 2
    x0 = -1 + (c + 0.5) * h
    hfm = Eikonal.dictIn({'model': 'Isotropic3', 'seeds': [[x0] * 3], 'exportValues': True,
                          'verbosity': 0})
    hfm.SetRect([[-1, 1]] * 3, dims=[n] * 3)
    out = hfm.Run()
    st = out['stats']['eikonal']
    u = out['values']
    log(f"isotropic3 {n}^3: passes {st['launches']}, cell-updates {st['cell_updates']}, "
        f"{st['gpu_time'] * 1e3:.0f} ms")
    assert st['converged']
    hf = float(np.float32(2.) / np.float32(n))
    ax = (torch.arange(n, device=u.device, dtype=torch.float64) - c) * hf
    for line in (u[:, c, c], u[c, :, c], u[c, c, :]):
        assert float((line.double() - ax.abs()).abs().max()) <= 1e-4
    worst = 0.
    for i in range(0, n, 65):                      # slabs of 65 planes
        sl = u[i:i + 65].double()
        xi = ax[i:i + 65].view(-1, 1, 1)
        ex = torch.sqrt(xi * xi + ax.view(1, -1, 1) ** 2 + ax.view(1, 1, -1) ** 2)
        worst = max(worst, float((sl - ex).abs().max()))
        # symmetry under exchanging the last two axes
        assert float((sl - sl.transpose(1, 2)).abs().max()) <= 1e-4
    # the oracle's first-order error on a 65^3 grid bounds the finer grid's
    m = 65
    hm = 2. / m
    P = O.make_problem('Isotropic', (m, m, m), [hm] * 3, seeds_idx=[[m // 2] * 3])
    xm = (np.arange(m) - m // 2) * hm
    Xm = np.stack(np.meshgrid(xm, xm, xm, indexing='ij'))
    err_m = float(np.abs(P.solve_fm() - np.sqrt((Xm ** 2).sum(0))).max())
    log(f"  max error {worst:.3e} (oracle 65^3: {err_m:.3e})")
    assert worst <= err_m
