"""CPU restatement of the geodesic backtracking of the reference's GPU
eikonal solver.  Test infrastructure only: imported by tests/, never by the
product path.

Follows (paths relative to the AGD source tree):
  * _GetGeodesics.py:50-67, inf_convolution.py:24-93 and
    Kernel_InfConvolution.h: the capped distance to the seeds `eucl`
    (0 on the seeds, `niter` min-plus passes with the integer kernel
    round(mult |offset|) over the 3^d neighbourhood, saturation at 255,
    values above the target tolerance set to 255);
  * Kernel_GeodesicODE.h:116-260 (NormalizedFlow): multilinear interpolation
    of the flow at the corners of the current cell, corners at or above the
    causality threshold left out, Euclidean normalisation, stopping codes
    AtSeed / InWall / VanishingFlow;
  * Kernel_GeodesicODE.h:265-413 (GeodesicODE): midpoint steps, the
    Stationnary and PastSeed histories (including the reference's indexing of
    the PastSeed history by a boolean, i.e. slot 0 or 1);
  * _GetGeodesics.py:86-131: launches of `typical_len` points until every
    geodesic stopped or max_len points were produced;
  * charts (_GetGeodesics.py:66-71, Kernel_GeodesicODE.h:317-320, 406,
    GeodesicODE_Opt.h:43-99): a step taken from a cell marked `chart_jump`
    keeps the previous distance to the seeds for the tests, then the point's
    last coordinates are replaced by the multilinear interpolation of the
    chart mapping at their cell's corners (clamped to the grid).

Arrays are cell-major (not tiled), FP64 arithmetic.
"""
import itertools

import numpy as np

STOP_NAMES = ('Continue', 'AtSeed', 'InWall', 'Stationnary', 'PastSeed', 'VanishingFlow')


def eucl_distance(seeds, niter, mult=5, bound=30, periodic=None):
    """uint8 distance to the seed cells (bool array), in 1/mult pixels."""
    d = seeds.ndim
    periodic = (False,) * d if periodic is None else tuple(periodic)
    kern = {m: int(np.round(mult * np.sqrt(m))) for m in range(d + 1)}
    a = np.where(seeds, 0, 255).astype(np.int64)
    shape = np.array(seeds.shape)
    idx = np.indices(seeds.shape)
    for _ in range(niter):
        out = np.full_like(a, 255)
        for off in itertools.product((-1, 0, 1), repeat=d):
            z = [idx[k] + off[k] for k in range(d)]
            ok = np.ones(seeds.shape, bool)
            for k in range(d):
                if periodic[k]:
                    z[k] = np.mod(z[k], shape[k])
                else:
                    ok &= (z[k] >= 0) & (z[k] < shape[k])
                    z[k] = np.clip(z[k], 0, shape[k] - 1)
            v = a[tuple(z)]
            b = kern[sum(o != 0 for o in off)]
            prod = np.where(b >= 255 - v, 255, v + b)
            out = np.where(ok, np.minimum(out, prod), out)
        a = out
    a[a > bound] = 255
    return a.astype(np.uint8)


class _Field:
    def __init__(self, u, flow, wsum, eucl, periodic):
        self.u, self.flow, self.wsum, self.eucl = u, flow, wsum, eucl
        self.shape = u.shape
        self.d = u.ndim
        self.periodic = (False,) * self.d if periodic is None else tuple(periodic)

    def index(self, y):
        """Grid.h InRange_per + Index_tot: wrapped position, or None outside."""
        z = []
        for k in range(self.d):
            v = int(y[k])
            if self.periodic[k]:
                v %= self.shape[k]
            elif v < 0 or v >= self.shape[k]:
                return None
            z.append(v)
        return tuple(z)


class _Cell:
    def __init__(self, d):
        nc = 1 << d
        self.xq = [None] * d
        self.flow = np.zeros((nc, d))
        self.dist = np.full(nc, np.inf)
        self.thr = 0.
        self.nymin = None


def _normalized_flow(F, x, c, causality):
    d = F.d
    nc = 1 << d
    xq = [int(np.floor(v)) for v in x]
    fresh = xq != c.xq
    c.xq = xq
    dx = [x[k] - xq[k] for k in range(d)]
    w = np.ones(nc)
    for i in range(nc):
        for k in range(d):
            w[i] *= dx[k] if (i >> k) & 1 else 1. - dx[k]
    if fresh:
        ws = np.zeros(nc)
        pos = [None] * nc
        for i in range(nc):
            p = F.index([xq[k] + ((i >> k) & 1) for k in range(d)])
            pos[i] = p
            if p is None:
                c.dist[i] = np.inf
                continue
            c.dist[i] = F.u[p]
            ws[i] = F.wsum[p]
            c.flow[i] = F.flow[(slice(None),) + p]
        wthr = 0.5 / nc
        dmin, imin = np.inf, 0
        for i in range(nc):
            if w[i] < wthr:
                continue
            if c.dist[i] < dmin:
                dmin, imin = c.dist[i], i
        if dmin == np.inf:
            return 2, None
        c.nymin = pos[imin]
        if ws[imin] == 0.:
            return 1, None
        c.thr = dmin + causality / ws[imin]
    flow = np.zeros(d)
    for i in range(nc):
        if c.dist[i] >= c.thr:
            continue
        flow += w[i] * c.flow[i]
    nrm = np.sqrt(np.sum(flow ** 2))
    if not nrm > 0:
        return 5, None
    return 0, flow / nrm


EUCL_CHART = 254


def _chart_jump(F, x):
    """x with its last ndim_s coordinates mapped to the other chart."""
    m = F.chart
    ns = m.shape[0]
    nb = F.d - ns
    shape_s = m.shape[1:]
    xs = x[nb:]
    q = np.floor(xs).astype(np.int64)
    r = xs - q
    out = np.zeros(ns)
    vals = []
    for corner in itertools.product((0, 1), repeat=ns):
        w = 1.
        idx = []
        for i, e in enumerate(corner):
            w *= r[i] if e else 1. - r[i]
            idx.append(int(np.clip(q[i] + e, 0, shape_s[i] - 1)))
        v = m[(slice(None),) + tuple(idx)]
        out += w * v
        vals.append(v)
    vals = np.array(vals)
    var = float(np.sum(np.mean(vals ** 2, axis=0) - np.mean(vals, axis=0) ** 2))
    if var >= F.chart_var:
        idx = tuple(int(np.clip(q[i] + (1 if r[i] > 0.5 else 0), 0, shape_s[i] - 1)) for i in range(ns))
        out = m[(slice(None),) + idx]
    if np.all(np.isfinite(out)):
        x = x.copy()
        x[nb:] = out
    return x


def _ode(F, x0, max_len, step, causality, eucl_delay, nymin_delay, pastseed_delay):
    """One launch from x0: the points 1 .. len-1, the stop code."""
    hlen = 1 + max(eucl_delay, nymin_delay)
    eucl_p = [255] * hlen
    nymin_p = [None] * hlen
    x = np.array(x0, dtype=np.float64)
    c = _Cell(F.d)
    pts = []
    stop = 0
    ln = 1
    while ln < max_len:
        l = ln % hlen
        xp = x.copy()
        stop, flow = _normalized_flow(F, x, c, causality)
        if stop:
            break
        nymin_p[l] = c.nymin
        eucl_p[l] = int(F.eucl[c.nymin])
        chart = F.chart is not None and eucl_p[l] == EUCL_CHART
        if chart:
            eucl_p[l] = eucl_p[(l - 1 + hlen) % hlen]
        if c.nymin == nymin_p[(l - nymin_delay + hlen) % hlen]:
            stop = 3
            break
        ip = 1 if (((l - eucl_delay + hlen) % hlen) != 0 and ln >= pastseed_delay) else 0
        if eucl_p[l] > eucl_p[ip]:
            stop = 4
            break
        xm = 0.5 * step * flow + xp
        stop, flow = _normalized_flow(F, xm, c, causality)
        if stop:
            break
        x = step * flow + xp
        pts.append(x.copy())
        ln += 1
        if chart:
            x = _chart_jump(F, x)
    return np.array(pts).reshape(-1, F.d), stop


def geodesics(u, flow, wsum, seeds, tips, step=0.25, causality=4., eucl_delay=None,
              nymin_delay=None, max_len=None, typical_len=None, target_tol=6, pastseed_delay=0,
              periodic=None, chart_mapping=None, chart_jump=None, chart_var=np.inf):
    """Geodesics (continuous grid indices, (len, d) each) from `tips`
    (continuous indices) and their stopping criteria."""
    u = np.asarray(u, dtype=np.float64)
    d = u.ndim
    eucl_delay = int(np.sqrt(d) / step) if eucl_delay is None else eucl_delay
    nymin_delay = int(8. * np.sqrt(d) / step) if nymin_delay is None else nymin_delay
    smax = int(np.max(u.shape))
    max_len = int(max(40, 20 * smax / step)) if max_len is None else max_len
    typical_len = min(max_len, int(max(40, 0.5 * smax / step))) if typical_len is None else typical_len
    eucl = eucl_distance(np.asarray(seeds, bool), int(np.ceil(target_tol)), 5, int(5 * target_tol),
                         periodic)
    if chart_mapping is not None:
        eucl = np.where(eucl == EUCL_CHART, 255, eucl)
        eucl = np.where(np.broadcast_to(np.asarray(chart_jump, bool), eucl.shape), EUCL_CHART, eucl)
    F = _Field(u, np.asarray(flow, dtype=np.float64), np.asarray(wsum, dtype=np.float64), eucl, periodic)
    F.chart = None if chart_mapping is None else np.asarray(chart_mapping, dtype=np.float64)
    F.chart_var = chart_var
    out, stops = [], []
    for tip in np.asarray(tips, dtype=np.float64).reshape(-1, d):
        parts = [tip.reshape(1, -1)]
        stop = None
        delay = pastseed_delay
        for _ in range(int(np.ceil(max_len / typical_len))):
            pts, code = _ode(F, parts[-1][-1], typical_len, step, causality, eucl_delay, nymin_delay,
                             delay)
            parts.append(pts)
            delay = 0
            if code != 0:
                stop = STOP_NAMES[code]
                break
        out.append(np.concatenate(parts, axis=0))
        stops.append(stop if stop is not None else "Stopping criterion")
    return out, stops
