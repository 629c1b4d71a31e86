"""Batched solves: many independent eikonal problems in one device worklist.

    from adaptivegriddiscretizations_amd.Eikonal import batch
    outs = batch.RunBatch([hfmIn_0, hfmIn_1, ...])           # one hfmOut per problem
    best, table = batch.autotune([hfmIn_0, ...])               # tile shape / niter_i

The reference solves one grid per call, with a host round trip per grid pass
(_solvers.py:184-190); many small grids therefore leave the GPU mostly idle.
Here the tiles of all problems are concatenated into one tile space, each
tile knowing its problem's grid and constants, and a single worklist drives
every problem: one kernel launch per grid pass for the whole batch
(agd_eikonal_solve_batched).  Each problem's preprocessing and
postprocessing (geometry, seeds, flow, AD) is that of the single solve
(HFM_GPU.Interface); only the fixed-point iteration is shared.

Problems of a batch share the model, dimension, scalar type (FP32), solver
and kernel traits; each keeps its own grid, costs, seeds and tolerances.
Walls, periodic axes and precomputed schemes are not batched.
"""
import ctypes
import time

import numpy as np

from ._device import torch, ptr, stream_handle

# Autotuning results, keyed by (model, bucketed problem sizes).
_TUNED = {}


def _interfaces(hfmIns, traits):
    from . import HFM_GPU
    ifaces = []
    for hfm in hfmIns:
        h = hfm.copy()
        if traits is not None:
            h['traits'] = {**(h['traits'] if 'traits' in h else {}), **traits}
        h['precompute_scheme'] = False
        if 'verbosity' not in h:
            h['verbosity'] = 0
        it = HFM_GPU.Interface(h, batched=True)
        it.SetKernelTraits()
        it.SetGeometry()
        it.SetArgs()
        if it.wallDist_t is not None or any(it.periodic):
            raise ValueError("batched solves take no walls and no periodic axes")
        ifaces.append(it)
    first = ifaces[0]
    for it in ifaces[1:]:
        same = (it.model == first.model and it.float_t == first.float_t and it.solver == first.solver
                and it.shape_i == first.shape_i and it.niter_i == first.niter_i
                and it.fim_front_width == first.fim_front_width)
        if not same:
            raise ValueError("problems of a batch must share model, scalar type, solver and traits")
    return ifaces


class Batch:
    """Concatenated device buffers of several prepared problems."""

    def __init__(self, hfmIns, traits=None):
        from .. import _lib
        self._lib = _lib
        if len(hfmIns) == 0:
            raise ValueError("empty batch")
        self.ifaces = _interfaces(hfmIns, traits)
        it0 = self.ifaces[0]
        if it0.float_t != np.float32:
            raise ValueError("batched solves are FP32 only")
        self.sizes = [it.size_tot for it in self.ifaces]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        cat = lambda name: torch.cat([getattr(it, name) for it in self.ifaces])
        self.values_t = cat('values_t')
        self.values0_t = self.values_t.clone()
        self.rhs_t = cat('rhs_t')
        self.seeds_t = self._cat_bits('seeds_t')
        self.trigger_t = self._cat_bits('trigger_t')
        self.geom = None
        if it0.geom is not None:
            self.geom = torch.cat([it.geom for it in self.ifaces], dim=1).contiguous()
        Inst = _lib.BatchInstance
        self.inst = (Inst * len(self.ifaces))()
        for i, it in enumerate(self.ifaces):
            for k in range(3):
                self.inst[i].shape[k] = int(it.shape[k]) if k < it.ndim else 1
                self.inst[i].iso_weights[k] = float(it.iso_weights[k])
            self.inst[i].atol = it.atol
            self.inst[i].rtol = it.rtol
        self.stream = stream_handle()
        self.solver = _lib.Solver(self.stream)  # worklist workspace, cached pass graph
        self.ncells = int(sum(int(np.prod(it.shape)) for it in self.ifaces))

    def _cat_bits(self, name):
        """Concatenation of the problems' bit arrays (1 bit per cell, 32 to a
        word): word by word when every size_tot is a multiple of 32, else
        bit by bit (tiles of 16 cells)."""
        words = [getattr(it, name) for it in self.ifaces]
        if all(n % 32 == 0 for n in self.sizes):
            return torch.cat(words)
        sh = torch.arange(32, device=words[0].device, dtype=torch.int64)
        bits = torch.cat([((w.to(torch.int64).unsqueeze(-1) >> sh) & 1).reshape(-1)[:n]
                          for w, n in zip(words, self.sizes)])
        pad = (-bits.numel()) % 32
        bits = torch.cat([bits, bits.new_zeros(pad)]).reshape(-1, 32)
        packed = torch.sum(bits << sh, dim=1)                       # < 2^32
        return torch.where(packed >= 2 ** 31, packed - 2 ** 32, packed).to(torch.int32)

    def desc(self):
        d = self.ifaces[0]._desc()
        d.nitermax_o = max(it.nitermax_o for it in self.ifaces)
        return d

    def solve(self, reset=True, desc=None):
        """One batched solve (values reset to their initial state first)."""
        L = self._lib
        if reset:
            self.values_t.copy_(self.values0_t)
        io = L.EikonalIO()
        io.values = ptr(self.values_t)
        io.geom = ptr(self.geom)
        io.rhs = ptr(self.rhs_t)
        io.seeds = ptr(self.seeds_t)
        io.trigger = ptr(self.trigger_t)
        st = L.EikonalStats()
        d = self.desc() if desc is None else desc
        status = self.solver.solve_batched(d, len(self.ifaces), self.inst, io, st)
        if status == L.AGD_NOCONV and not all(it.raiseOnNonConvergence for it in self.ifaces):
            self.ifaces[0].Warn("batched solve did not converge")
        else:
            L.check(status)
        return st

    def scatter(self):
        """Copy each problem's slice of the batch values back to its own
        interface (for its postprocessing and outputs)."""
        for i, it in enumerate(self.ifaces):
            it.values_t.copy_(self.values_t[self.offsets[i]:self.offsets[i + 1]])
            it._values = None

    def refresh_values(self):
        """Re-gather the batch's initial values from the problems' own
        buffers (after they were changed in place, e.g. repaired starts)."""
        self.values_t.copy_(torch.cat([it.values_t for it in self.ifaces]))
        self.values0_t.copy_(self.values_t)

    def values(self, i):
        """Cell-major values of problem i."""
        self.ifaces[i].values_t.copy_(self.values_t[self.offsets[i]:self.offsets[i + 1]])
        self.ifaces[i]._values = None
        return self.ifaces[i].values


def RunBatch(hfmIns, traits=None, tune=False):
    """Solve the dictIn problems hfmIns in one batched worklist and return
    their hfmOut dicts (postprocessing per problem, as Interface.Run)."""
    if tune and traits is None:
        traits = _TUNED.get(_key(hfmIns)) or autotune(hfmIns)[0]
    b = Batch(hfmIns, traits)
    torch.cuda.synchronize()
    t0 = time.time()
    st = b.solve(reset=False)
    torch.cuda.synchronize()
    wall = time.time() - t0
    b.scatter()
    outs = []
    for it in b.ifaces:
        it.stats['eikonal'] = {**st.as_dict(), 'time': wall, 'gpu_time': st.gpu_ms * 1e-3, 'batched': True}
        it._values = None
        it.PostProcess()
        it.SolveAD()
        it.FinalCheck()
        outs.append(it.hfmOut)
    return outs


def _key(hfmIns):
    model = hfmIns[0]['model']
    sizes = sorted(int(np.prod(h.shape)) for h in hfmIns)
    bucket = tuple(int(np.log2(max(s, 1))) for s in (sizes[0], sizes[len(sizes) // 2], sizes[-1]))
    return (model, len(hfmIns), bucket)


def candidate_traits(model):
    """Tile shapes and in-tile sweep counts tried by the autotuner."""
    if model.endswith('2'):
        return [{'shape_i': s, 'niter_i': n}
                for s in ((4, 4), (4, 8), (8, 4), (8, 8), (16, 16), (8, 16)) for n in (4, 8, 16)]
    return [{'shape_i': s, 'niter_i': n} for s in ((4, 4, 4), (8, 8, 4), (4, 4, 8)) for n in (2, 4, 8)]


def autotune(hfmIns, candidates=None, repeats=2):
    """Time the batched solve of hfmIns for each candidate trait set and
    keep the fastest (cached for batches of the same model and size
    profile).  Returns (best traits, [(traits, ms, passes), ...])."""
    from .. import _lib
    model = hfmIns[0]['model']
    table = []
    for tr in (candidates or candidate_traits(model)):
        if not _lib.lib.agd_batch_tile_supported(
                {'Isotropic': _lib.ISOTROPIC, 'Diagonal': _lib.DIAGONAL, 'Riemann': _lib.RIEMANN,
                 'Rander': _lib.RANDER}[model[:-1]], int(model[-1]), _lib.i32(tr['shape_i'])):
            continue
        b = Batch(hfmIns, tr)
        b.solve()  # warm-up (graph capture, workspace growth)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            st = b.solve()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / repeats * 1e3
        table.append((tr, ms, int(st.launches)))
        del b
    best = min(table, key=lambda r: r[1])[0]
    _TUNED[_key(hfmIns)] = best
    return best, table
