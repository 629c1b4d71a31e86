"""Host-side preparation of seed spreading and source factorization for the
GPU eikonal solver.

Reference: agd/Eikonal/HFM_CUDA/_SetArgs.py:36-78 (seedRadius spreading,
factor metric), _Interface.py:202-211 (CostMetric: the metric with the grid
steps and the cost folded in, interpolated at order 1), _Kernel.py:301-308
(factor_origin, factor_radius2, factor_metric constants).

Everything here works on a handful of points (the cells of one seed's
neighbourhood and the corners of their interpolation stencils): the fields
are gathered at those cells where they live (device tensors stay on the
device, only the gathered values travel) and the arithmetic is float64 numpy.
"""
import numpy as np

from ._device import torch

DIAGONAL_MODELS = ('Isotropic', 'Diagonal')


def gather_cells(arr, core, cells):
    """Values of a field arr[*core, *shape] (or arr[*core], constant over
    the grid) at the integer cells[k, d]: float64 array [k, *core]."""
    cells = np.asarray(cells, dtype=np.int64)
    k = len(cells)
    if isinstance(arr, torch.Tensor):
        t = arr.detach()
        if tuple(t.shape) == tuple(core):
            v = t.to(torch.float64).cpu().numpy()
            return np.broadcast_to(v, (k,) + tuple(core)).copy()
        idx = tuple(torch.as_tensor(cells[:, c], device=t.device) for c in range(cells.shape[1]))
        v = t[(Ellipsis,) + idx].to(torch.float64).cpu().numpy()
    else:
        a = np.asarray(arr, dtype=np.float64)
        if a.shape == tuple(core):
            return np.broadcast_to(a, (k,) + tuple(core)).copy()
        v = a[(Ellipsis,) + tuple(cells.T)]
    return np.moveaxis(v, -1, 0) if core else v


def interpolation_stencil(pts, shape, periodic):
    """Order-1 (multilinear) interpolation at the points pts[k, d] in grid
    units: corner cells [2^d, k, d] and weights [2^d, k].  On a non-periodic
    axis the point is clamped onto [0, n - 1]; periodic axes wrap."""
    pts = np.asarray(pts, dtype=np.float64)
    k, d = pts.shape
    base = np.empty((k, d), dtype=np.int64)
    frac = np.empty((k, d))
    for c in range(d):
        n = int(shape[c])
        x = pts[:, c] if periodic[c] else np.clip(pts[:, c], 0., n - 1.)
        b = np.floor(x) if periodic[c] else np.minimum(np.floor(x), max(n - 2, 0))
        base[:, c] = b.astype(np.int64)
        frac[:, c] = x - b
    corners, weights = [], []
    for m in range(1 << d):
        hi = np.array([(m >> c) & 1 for c in range(d)], dtype=np.int64)
        cell = base + hi
        for c in range(d):
            n = int(shape[c])
            cell[:, c] = np.mod(cell[:, c], n) if periodic[c] else np.minimum(cell[:, c], n - 1)
        corners.append(cell)
        weights.append(np.prod(np.where(hi == 1, frac, 1. - frac), axis=1))
    return np.stack(corners), np.stack(weights)


def tti_norm(l, q3, A, v, niter=6):
    """The TTI norm N(v) = N0(A v) at rows v[k, d]: N0(w) is the largest <p, w>
    over the dual ball f(p) <= 0, f(p) = l0 P0 + l1 P1 + (q00 P0^2 +
    2 q10 P0 P1 + q11 P1^2) / 2 - 1 with P0 = X^2 [+ Y^2] and P1 = Z^2 the
    squares of p's lateral and last coordinates (Seismic/tti.py:64-69).  The
    maximiser is found by sequential quadratic programming from p = 0
    (Seismic/implicit_base.py:83-95, 140-179: niter steps
    p += M (k w - grad f), M = hess f^-1, k = sqrt((<grad f, M grad f> - 2 f)
    / <w, M w>)); N0(w) = <p, w>.  l[k, 2], q3[k, 3] = (q00, q10, q11),
    A[k, d, d] (each also broadcast from one row)."""
    v = np.atleast_2d(np.asarray(v, dtype=np.float64))
    k, d = v.shape
    l = np.broadcast_to(np.asarray(l, dtype=np.float64).reshape(-1, 2), (k, 2))
    q3 = np.broadcast_to(np.asarray(q3, dtype=np.float64).reshape(-1, 3), (k, 3))
    A = np.broadcast_to(np.asarray(A, dtype=np.float64).reshape(-1, d, d), (k, d, d))
    w = np.einsum('kij,kj->ki', A, v)
    zero = np.all(w == 0., axis=1)
    w = np.where(zero[:, None], 1., w)
    cls = np.array([0] * (d - 1) + [1])                  # lateral (0) or last (1) coordinate
    qm = np.stack([np.stack([q3[:, 0], q3[:, 1]], -1), np.stack([q3[:, 1], q3[:, 2]], -1)], -2)
    qc = qm[:, cls][:, :, cls]                            # q between the coordinates' classes
    p = np.zeros_like(w)
    for _ in range(niter):
        P = np.stack([np.sum(p[:, :d - 1] ** 2, axis=1), p[:, d - 1] ** 2], -1)
        g = l + np.einsum('kij,kj->ki', qm, P)            # df / dP
        f = np.sum(l * P, axis=1) + 0.5 * np.einsum('ki,kij,kj->k', P, qm, P) - 1.
        gc = g[:, cls]
        V = 2. * p * gc                                   # grad f
        H = 2. * gc[:, :, None] * np.eye(d) + 4. * qc * p[:, :, None] * p[:, None, :]
        M = np.linalg.inv(H)
        kk = np.sqrt((np.einsum('ki,kij,kj->k', V, M, V) - 2. * f) / np.einsum('ki,kij,kj->k', w, M, w))
        p = p + np.einsum('kij,kj->ki', M, kk[:, None] * w - V)
    return np.where(zero, 0., np.sum(p * w, axis=1))


class CostMetric:
    """The metric of a solve in grid units at a few points: Isotropic costs
    h_i c, Diagonal costs_i h_i c, Riemann m_ij h_i h_j c^2, Rander
    (m_ij h_i h_j c^2, w_i h_i c), TTI (l, q, A diag(h) c), with c the cost
    (Metrics.*.with_costs(h) followed by with_cost(cost)), interpolated at
    order 1.

    diag[k, d] (diagonal models), m[k, d, d] and w[k, d] (or None), or for
    TTI the tuple (l[k, 2], q3[k, 3], A[k, d, d]); metric is then the tuple
    (linear, quadratic, inverse_transformation or None)."""

    def __init__(self, model, h, cost, metric=None, drift=None):
        self.model = model
        self.h = np.asarray(h, dtype=np.float64)
        self.cost = cost
        self.metric = metric   # Diagonal: costs[d, ...]; Riemann/Rander: m[d, d, ...]
        self.drift = drift     # Rander: w[d, ...]

    def _at_cells(self, cells):
        d = len(self.h)
        c = gather_cells(self.cost, (), cells)
        if self.model == 'TTI':
            lin, quad, A = self.metric
            l = gather_cells(lin, (2,), cells)
            q = gather_cells(quad, (2, 2), cells)
            q3 = np.stack([q[:, 0, 0], q[:, 1, 0], q[:, 1, 1]], axis=1)
            Am = (np.broadcast_to(np.eye(d), (len(cells), d, d)) if A is None
                  else gather_cells(A, (d, d), cells))
            return (l, q3, Am * self.h[None, None, :] * c[:, None, None]), None
        if self.model in DIAGONAL_MODELS:
            costs = np.ones((len(cells), d)) if self.model == 'Isotropic' \
                else gather_cells(self.metric, (d,), cells)
            return costs * self.h * c[:, None], None
        m = gather_cells(self.metric, (d, d), cells) * np.outer(self.h, self.h) * (c * c)[:, None, None]
        w = None
        if self.drift is not None:
            w = gather_cells(self.drift, (d,), cells) * self.h * c[:, None]
        return m, w

    def at(self, pts, shape, periodic):
        """Returns (a, w): a = costs [k, d] or m [k, d, d]; w the Rander drift [k, d] or None."""
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, len(self.h))
        corners, weights = interpolation_stencil(pts, shape, periodic)
        nc, k, d = corners.shape
        a, w = self._at_cells(corners.reshape(-1, d))

        def interp(f):
            f = f.reshape((nc, k) + f.shape[1:])
            return np.sum(f * weights.reshape((nc, k) + (1,) * (f.ndim - 2)), axis=0)

        a = tuple(interp(f) for f in a) if isinstance(a, tuple) else interp(a)
        if w is not None:
            w = interp(w)
        return a, w

    def norm(self, a, w, v):
        """|v|: sqrt(sum (c_i v_i)^2), or sqrt(v m v) (+ <w, v> for Rander)."""
        v = np.asarray(v, dtype=np.float64)
        if self.model == 'TTI':
            return tti_norm(a[0], a[1], a[2], v.reshape(-1, v.shape[-1]))
        if self.model in DIAGONAL_MODELS:
            return np.sqrt(np.sum((a * v) ** 2, axis=-1))
        out = np.sqrt(np.einsum('...i,...ij,...j->...', v, a, v))
        if w is not None:
            out = out + np.sum(w * v, axis=-1)
        return out

    def factor_metric(self, a):
        """The factorization metric of one point (_Kernel.py:304-308): the
        squared costs (diagonal models), [l0, l1, q00, q10, q11, A row-major]
        (TTI), else the lower triangle (m00, m10, m11, m20, m21, m22) of m;
        the drift is left out."""
        if self.model == 'TTI':
            return [float(x) for x in np.concatenate([np.asarray(f, dtype=np.float64).reshape(-1)
                                                      for f in a])]
        a = np.asarray(a, dtype=np.float64)
        if self.model in DIAGONAL_MODELS:
            return list((a.reshape(-1)) ** 2)
        a = a.reshape(a.shape[-2:])
        d = a.shape[0]
        return [a[i, j] for i in range(d) for j in range(i + 1)]


def spread_cells(seed, radius, shape, periodic):
    """The cells replacing one seed (_SetArgs.py:46-66): integer points at
    Euclidean distance < radius (grid units) from the seed, inside the
    domain along the non-periodic axes.  Returns (cells[k, d], diff[k, d])
    with cells wrapped on periodic axes and diff = cell - seed unwrapped."""
    seed = np.asarray(seed, dtype=np.float64).reshape(-1)
    d = len(seed)
    ranges = [np.arange(int(np.floor(s - radius)), int(np.ceil(s + radius)) + 1) for s in seed]
    pts = np.stack(np.meshgrid(*ranges, indexing='ij'), axis=-1).reshape(-1, d)
    keep = np.sqrt(np.sum((pts - seed) ** 2, axis=1)) < radius
    for c in range(d):
        if not periodic[c]:
            keep &= (pts[:, c] >= -0.5) & (pts[:, c] < shape[c] - 0.5)
    pts = pts[keep]
    diff = pts - seed
    cells = pts.copy()
    for c in range(d):
        if periodic[c]:
            cells[:, c] = np.mod(cells[:, c], shape[c])
    return cells.astype(np.int64), diff


def spread_seed(cm, seed, radius, shape, periodic, value):
    """Seed spreading (_SetArgs.py:46-75): cells of spread_cells, valued
    value + (|x - seed|_{M(seed)} + |x - seed|_{M(x)}) / 2 with the cost
    metric M.  Returns (cells[k, d], values[k])."""
    cells, diff = spread_cells(seed, radius, shape, periodic)
    a0, w0 = cm.at(np.asarray(seed, dtype=np.float64)[None], shape, periodic)
    a1, w1 = cm.at(cells, shape, periodic)
    values = value + 0.5 * (cm.norm(a0, w0, diff) + cm.norm(a1, w1, diff))
    return cells, values
