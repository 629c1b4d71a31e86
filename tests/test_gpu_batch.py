"""Batched solves: every problem of a batch reaches the fixed point of its
own unbatched solve (same finite mask, values within the solver tolerance),
for each solver and for the autotuned traits."""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import Eikonal, workloads
from gpu_helpers import log, oracle_problem, rel_err

pytestmark = pytest.mark.gpu


def iso_problem(rng, n0, n1, nseeds):
    cost = (1 + 0.5 * np.tanh(rng.normal(size=(n0, n1)))).astype(np.float32)
    hfm = Eikonal.dictIn({'model': 'Isotropic2', 'cost': cost, 'verbosity': 0, 'exportValues': True,
                          'seeds': rng.uniform(-0.9, 0.9, (nseeds, 2))})
    hfm.SetRect([[-1, 1], [-1, 1]], dims=[n0, n1])
    return hfm


def riemann_problem(n, seed):
    wl = workloads.riemann2_aniso(n=n, seed=seed, sigma=3., amax=6., nseeds=2)
    return wl.dict_in(exportValues=True)


def check_against_single(hfms, outs, tol=2e-5):
    for i, (hfm, out) in enumerate(zip(hfms, outs)):
        ref = hfm.Run()['values'].cpu().numpy().astype(np.float64)
        u = out['values'].cpu().numpy().astype(np.float64)
        err = rel_err(u, ref)
        log(f"  problem {i} {hfm.shape}: rel diff to the single solve {err:.2e}")
        assert err <= tol


@pytest.mark.parametrize("solver", ["FIM", "AGSI", "global_iteration"])
def test_isotropic_batch_matches_single_solves(solver):
    from adaptivegriddiscretizations_amd.Eikonal import batch
    rng = np.random.default_rng(7)
    shapes = [(70, 90), (128, 128), (33, 200), (64, 64), (150, 77), (96, 130)]
    hfms = [iso_problem(rng, a, b, 1 + i % 3) for i, (a, b) in enumerate(shapes)]
    for h in hfms:
        h['solver'] = solver
    outs = batch.RunBatch(hfms)
    log(f"batch of {len(hfms)} ({solver}): passes {outs[0]['stats']['eikonal']['launches']}")
    check_against_single(hfms, outs)


def test_riemann_batch_matches_single_solves_and_oracle(O):
    from adaptivegriddiscretizations_amd.Eikonal import batch
    hfms = [riemann_problem(n, 20 + n) for n in (40, 72, 100)]
    outs = batch.RunBatch(hfms)
    check_against_single(hfms, outs)
    wl = workloads.riemann2_aniso(n=72, seed=92, sigma=3., amax=6., nseeds=2)
    P, _ = oracle_problem(O, wl)
    u = outs[1]['values'].cpu().numpy().astype(np.float64)
    assert rel_err(u, P.solve_fm()) <= 1e-4


def test_autotune_picks_a_candidate_and_is_reused():
    from adaptivegriddiscretizations_amd.Eikonal import batch
    rng = np.random.default_rng(8)
    hfms = [iso_problem(rng, 64 + 16 * i, 96, 2) for i in range(8)]
    best, table = batch.autotune(hfms, repeats=1)
    for tr, ms, passes in table:
        log(f"  {tr}: {ms:.2f} ms, {passes} passes")
    assert best in [t for t, _, _ in table]
    assert best == min(table, key=lambda r: r[1])[0]
    # the candidates span 4x4 to 16x16 tiles (tiles below a wave included)
    shapes = {tuple(t['shape_i']) for t, _, _ in table}
    assert {(4, 4), (4, 8), (8, 4), (8, 8), (16, 16)} <= shapes
    outs = batch.RunBatch(hfms, tune=True)
    check_against_single(hfms[:2], outs[:2])


@pytest.mark.parametrize("shape_i", [(4, 4), (4, 8), (8, 4)])
@pytest.mark.parametrize("solver", ["FIM", "AGSI"])
def test_subwave_tiles_batch_matches_single_solves(solver, shape_i):
    """Tiles of 16 and 32 cells, packed 4 or 2 to a wave: every problem of
    the batch reaches the fixed point of its single solve (8x8 tiles, the
    same scheme), including odd tile counts whose seed bits do not start on
    a word boundary, and an empty lane group at the end of every pass."""
    from adaptivegriddiscretizations_amd.Eikonal import batch
    rng = np.random.default_rng(11)
    shapes = [(70, 90), (33, 200), (17, 45), (64, 64), (150, 77), (5, 9)]
    hfms = [iso_problem(rng, a, b, 1 + i % 3) for i, (a, b) in enumerate(shapes)]
    for h in hfms:
        h['solver'] = solver
    outs = batch.RunBatch(hfms, traits={'shape_i': shape_i, 'niter_i': 8})
    log(f"batch of {len(hfms)} ({solver}, tiles {shape_i}): passes "
        f"{outs[0]['stats']['eikonal']['launches']}")
    check_against_single(hfms, outs)


def test_subwave_tiles_are_refused_outside_batches():
    rng = np.random.default_rng(12)
    hfm = iso_problem(rng, 40, 40, 1)
    hfm['traits'] = {'shape_i': (4, 4)}
    with pytest.raises(ValueError, match="tile shape"):
        hfm.Run()
