"""Re-entrancy of the C ABI: two solves through two solver handles, on two
streams of one device, issued at the same time from two host threads
(ctypes releases the interpreter lock during the calls), each checked
against the oracle's FP64 fast marching of its own problem; and a handle
reused across solves (the cached pass graph is replayed) giving the
oracle's answer each time.

Bar: max|u - u_oracle| <= 1e-4 max u (FP32), identical finite masks.
"""
import threading

import numpy as np
import pytest

from adaptivegriddiscretizations_amd import _lib, workloads
from gpu_helpers import log, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

TOL32 = 1e-4


def _prepare(wl):
    from adaptivegriddiscretizations_amd.Eikonal import HFM_GPU
    iface = HFM_GPU.Interface(wl.dict_in())
    iface.SetKernelTraits()
    iface.SetGeometry()
    iface.SetArgs()
    return iface


def test_two_handles_on_two_streams_at_once(O):
    import torch
    wls = [workloads.riemann2_aniso(n=240, seed=21, sigma=5., amax=8.),
           workloads.isotropic2_bumps(n=320, seed=4)]
    ifaces = [_prepare(wl) for wl in wls]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    handles = [_lib.Solver(s.cuda_stream, it._desc()) for s, it in zip(streams, ifaces)]
    for it in ifaces:
        it.values_t.fill_(float('inf'))
    torch.cuda.synchronize()
    status = [None, None]
    stats = [_lib.EikonalStats(), _lib.EikonalStats()]
    start = threading.Barrier(2)

    def run(i):
        start.wait()
        status[i] = handles[i].solve(ifaces[i]._desc(), ifaces[i]._io(), stats[i])

    threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    log("two handles, two streams, two host threads")
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
        assert not t.is_alive(), "a concurrent solve did not return"
    torch.cuda.synchronize()
    for i, (wl, it) in enumerate(zip(wls, ifaces)):
        _lib.check(status[i])
        assert stats[i].converged == 1
        it._values = None
        u = it.values.cpu().numpy().astype(np.float64)
        P, _ = oracle_problem(O, wl)
        err = rel_err(u, P.solve_fm())
        log(f"  {wl.name} {wl.dims}: passes {stats[i].niter_o}, rel err {err:.2e}")
        assert err <= TOL32
    for h in handles:
        h.close()


def test_handle_reuse_replays_graph(O):
    """Three solves through one handle (the second and third replay the
    pass graph captured by the first); each reaches the oracle's fixed
    point, and a solve after a change of arguments (other tolerance)
    re-captures and still does."""
    import torch
    wl = workloads.riemann2_aniso(n=160, seed=23, sigma=5., amax=6.)
    it = _prepare(wl)
    P, _ = oracle_problem(O, wl)
    ref = P.solve_fm()
    h = it.solver_handle()
    for k, rtol in enumerate([None, None, 1e-7]):
        it.values_t.fill_(float('inf'))
        it._values = None
        d = it._desc()
        if rtol is not None:
            d.rtol = rtol
            d.atol = rtol
        st = _lib.EikonalStats()
        _lib.check(h.solve(d, it._io(), st))
        torch.cuda.synchronize()
        err = rel_err(it.values.cpu().numpy().astype(np.float64), ref)
        log(f"solve {k}: passes {st.niter_o}, rel err {err:.2e}")
        assert st.converged == 1 and err <= TOL32
