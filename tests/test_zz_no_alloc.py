"""The allocation-free solve (agd_eikonal_solve_noalloc, worklist in the
library's fixed device region) against the allocating entry point
(agd_eikonal_solve) on the same inputs, through the C ABI: two calls in a
row, the second on a smaller grid, so that the region's contents left by
the first call must not leak into the second.

Bar: the same discrete fixed point.  Within a pass, tiles visited at the
same time may read each other's values before or after their update, so
FP32 values of two solves agree to the stopping tolerance, not bit for bit:
max |u_noalloc - u_alloc| <= 1e-5 max u, identical finite masks, both
converged; the FP64 solve agrees to 1e-12 max u.
"""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import _lib, workloads
from gpu_helpers import log


def _prepare(wl, **extra):
    from adaptivegriddiscretizations_amd.Eikonal import HFM_GPU
    iface = HFM_GPU.Interface(wl.dict_in(**extra))
    iface.SetKernelTraits()
    iface.SetGeometry()
    iface.SetArgs()
    return iface


def _solve(iface, entry):
    import torch
    iface.values_t.fill_(float('inf'))
    iface._values = None
    st = _lib.EikonalStats()
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(entry(iface._desc(), iface._io(), st, stream))
    torch.cuda.synchronize()
    return iface.values.cpu().numpy().astype(np.float64), st


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_noalloc_matches_allocating_solve(dtype, tol):
    cases = [workloads.riemann2_aniso(n=300, seed=11, sigma=6., amax=8.),
             workloads.riemann2_aniso(n=96, seed=12, sigma=4., amax=5.)]
    for wl in cases:   # second call: smaller input
        log(f"noalloc vs alloc: riemann2 {wl.dims} {np.dtype(dtype).name}")
        extra = {'traits': {'Scalar': np.float64}} if dtype == np.float64 else {}
        iface = _prepare(wl, **extra)
        iface.atol = iface.rtol = (1e-6 if dtype == np.float32 else 1e-13)
        u_na, st_na = _solve(iface, _lib.lib.agd_eikonal_solve_noalloc)
        u_al, st_al = _solve(iface, _lib.lib.agd_eikonal_solve)
        assert st_na.converged == 1 and st_al.converged == 1
        fin = np.isfinite(u_al)
        assert np.array_equal(fin, np.isfinite(u_na))
        err = np.max(np.abs(u_na[fin] - u_al[fin])) / np.max(np.abs(u_al[fin]))
        log(f"  passes {st_na.niter_o} / {st_al.niter_o}, rel diff {err:.2e}")
        assert err <= tol


def test_noalloc_refuses_grids_beyond_region():
    """A grid whose worklist exceeds the fixed region is refused with
    AGD_EINVAL before any device call (runs without a GPU)."""
    d = _lib.EikonalDesc()
    d.model, d.ndim, d.scalar, d.solver = _lib.ISOTROPIC, 3, _lib.F32, _lib.FIM
    for k in range(3):
        d.shape[k] = 400
        d.shape_i[k] = 4
    d.niter_i, d.fim_front_width, d.nitermax_o, d.check_interval = 4, 6, 2000, 18
    io = _lib.EikonalIO()
    assert _lib.lib.agd_eikonal_solve_noalloc(d, io, None, None) == _lib.AGD_EINVAL
    assert b"fixed region" in _lib.lib.agd_last_error()
