"""bound_active_blocks on the GPU (_solvers.py:145-181,195-224,
Propagation.h:53-99): freezing tiles above a moving threshold changes the
schedule only, so the AGSI solve reaches the oracle's fixed point; the
option keeps the reference's restrictions (AGSI only)."""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import workloads
from gpu_helpers import log, oracle_problem, rel_err, run_gpu

pytestmark = pytest.mark.gpu
TOL32 = 1e-4


@pytest.mark.parametrize("bound", [True, 16, 200])
def test_riemann2_bound_active_blocks_matches_oracle(O, bound):
    wl = workloads.riemann2_aniso(n=160, seed=5, sigma=4., amax=6.)
    u, out = run_gpu(wl, bound_active_blocks=bound, solver='AGSI')
    u0, out0 = run_gpu(wl, solver='AGSI')
    P, _ = oracle_problem(O, wl)
    ref = P.solve_fm()
    err = rel_err(u, ref)
    st, st0 = out['stats']['eikonal'], out0['stats']['eikonal']
    log(f"riemann2 160^2 bound={bound}: rel err {err:.2e}; cell-updates {st['cell_updates']} "
        f"(plain AGSI {st0['cell_updates']}), passes {st['launches']} ({st0['launches']})")
    assert err <= TOL32
    assert np.array_equal(np.isfinite(u), np.isfinite(ref))


def test_isotropic2_bound_active_blocks_matches_oracle(O):
    wl = workloads.isotropic2_bumps(n=256, seed=4)
    u, out = run_gpu(wl, bound_active_blocks=True)
    P, _ = oracle_problem(O, wl)
    err = rel_err(u, P.solve_fm())
    log(f"isotropic2 256^2 bound=True: rel err {err:.2e}, cell-updates {out['stats']['eikonal']['cell_updates']}")
    assert err <= TOL32


def test_bound_active_blocks_solver_restrictions():
    wl = workloads.isotropic2_bumps(n=64, seed=1)
    for solver in ('FIM', 'global_iteration'):
        with pytest.raises(ValueError, match="bound_active_blocks"):
            run_gpu(wl, bound_active_blocks=True, solver=solver)
