"""Shared helpers of the GPU parity tests: build the same problem for the
HIP path (through Eikonal.dictIn) and for the CPU oracle."""
import sys

import numpy as np


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_problem(O, wl, walls=None, seed_values=None, cast=True, periodic=None):
    """Oracle problem matching workload `wl` (FP64 arithmetic).  With cast,
    geometry and rhs are rounded to float32 as the FP32 device path stores
    them; the FP64 device path keeps them in double."""
    sides = np.asarray(wl.sides, dtype=np.float64)
    h = wl.h.astype(np.float32).astype(np.float64)  # grid scales are float32 in GPU mode
    idx = O.seed_indices(wl.seeds, sides[:, 0], h)
    model = wl.model[:-1]
    kw = {}
    if model in ('Riemann', 'Rander'):
        kw['metric'] = wl.metric.astype(np.float64)
    if model == 'Rander':
        kw['drift'] = wl.extra['drift'].astype(np.float64)
    if model == 'Diagonal':
        kw['metric'] = wl.extra['costs'].astype(np.float64)
    cost = 1. if wl.cost is None else wl.cost.astype(np.float64)
    P = O.make_problem(model, wl.dims, h, cost=cost, seeds_idx=idx, seed_values=seed_values,
                       walls=walls, periodic=periodic, **kw)
    return (P.cast(np.float32) if cast else P), idx


def run_gpu(wl, **extra):
    hfmIn = wl.dict_in(exportValues=True, **extra)
    out = hfmIn.Run()
    return out['values'].cpu().numpy().astype(np.float64), out


def rel_err(u, ref):
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(u)), "finite masks differ"
    return np.max(np.abs(u[fin] - ref[fin])) / np.max(np.abs(ref[fin]))
