"""Python side of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads liboracle.so (agd_oracle.cpp, built by the Makefile next to this
file) and restates, in numpy FP64, the reference's host preprocessing of the
GPU eikonal path:
  * grid geometry: the box arithmetic of SetRect in float32, cell centres
    origin + (i+1/2) h, continuous index (x - origin)/h - 1/2, seeds rounded
    half to even (agd/Eikonal/DictIn.py:97-107,404-453,
    HFM_CUDA/_SetArgs.py:44-45);
  * the tiled (bi-level) layout, as a checker of the device's
    agd_block_expand / agd_block_squeeze (FiniteDifferences.py:164-222);
  * metric preparation: D = (diag(h) M diag(h))^-1 flattened lower-triangular
    for Riemann, [flat(inv(h m h)), h w] for Rander, (c h)^-2 for Diagonal,
    h^-2 for Isotropic (HFM_CUDA/_SetGeometry.py:127-142,
    HFM_CUDA/_Kernel.py:316-317, Metrics/riemann.py:38-73,
    Metrics/rander.py:61-68, Metrics/misc.py:8-15);
  * rhs = cost with rhs[seed] = seed value, walls -> rhs = inf and fixed
    (_SetArgs.py:116-126);
  * seed spreading over a radius and the source factorization metric
    (_SetArgs.py:46-78, _Interface.py:202-211, _Kernel.py:301-308).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ISOTROPIC, DIAGONAL, RIEMANN, RANDER, TTI = 0, 1, 2, 3, 4
REEDSSHEPP2, REEDSSHEPPFWD2, ELASTICA2, DUBINS2 = 5, 6, 7, 8
MODEL_IDS = {'Isotropic': ISOTROPIC, 'Diagonal': DIAGONAL, 'Riemann': RIEMANN, 'Rander': RANDER,
             'TTI': TTI, 'ReedsShepp2': REEDSSHEPP2, 'ReedsSheppForward2': REEDSSHEPPFWD2,
             'Elastica2': ELASTICA2, 'Dubins2': DUBINS2}
NMIX = 7
# offsets of one sub-stencil of the curvature models (the forward ones included)
CURVATURE_NSYM = {REEDSSHEPP2: 6, REEDSSHEPPFWD2: 7, ELASTICA2: 30, DUBINS2: 6}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def omp_threads():
    """Threads of the OpenMP tile solver (OMP_NUM_THREADS or the host's cores)."""
    return int(lib().orc_omp_threads())


def _load():
    src = os.path.join(HERE, "agd_oracle.cpp")
    if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    lib.orc_version.restype = c_int
    lib.orc_selling.argtypes = [c_int, c_int64, P, P, P, c_int]
    lib.orc_hfm_update.argtypes = [c_int, c_int64, P, P, P, P, P, c_int]
    lib.orc_stencils.argtypes = [c_int, c_int, P, P, P, c_int, P, P, P]
    lib.orc_tti_diags.argtypes = [c_int64, P, P, P, P, c_int]
    lib.orc_solve_fm.argtypes = [c_int, c_int, P, P, P, P, P, P, P, P]
    lib.orc_solve_fm.restype = c_int64
    lib.orc_solve_fm_from.argtypes = [c_int, c_int, P, P, P, P, P, P, P, P, P]
    lib.orc_solve_fm_from.restype = c_int64
    lib.orc_solve_jacobi.argtypes = [c_int, c_int, P, P, P, P, P, P, P, c_double, c_double,
                                     c_int64, c_int, P]
    lib.orc_solve_jacobi.restype = c_int64
    lib.orc_flow.argtypes = [c_int, c_int, P, P, P, P, P, P, P, c_int, P, P, P, P, P]
    lib.orc_adjoint.argtypes = [c_int64, c_int, P, P, P, c_int, P, P]
    lib.orc_forward_ad.argtypes = [c_int64, c_int, P, P, P, c_int, P, P]
    lib.orc_set_order.argtypes = [c_int, c_double]
    lib.orc_set_factor.argtypes = [c_double, P, P]
    lib.orc_set_curvature_relax.argtypes = [c_double]
    lib.orc_fejer.argtypes = [c_int, c_int, P]
    lib.orc_set_curvature_convex.argtypes = [c_int]
    lib.orc_tti_factor.argtypes = [c_int, c_int, P, c_int64, P, P, P, P, P, P]
    lib.orc_wall_distance.argtypes = [c_int, P, P, P, c_int, P]
    lib.orc_visible.argtypes = [c_int, P, P, P, c_int64, P, P, P]
    lib.orc_omp_threads.restype = c_int
    lib.orc_solve_tiles_omp.argtypes = [c_int, c_int, P, P, P, P, P, P, P, c_int, c_int, c_double,
                                        c_double, c_int64, P, P]
    lib.orc_solve_tiles_omp.restype = c_int64
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- kernels
def selling(m, use_float=False):
    """Selling decomposition of matrices m[n, symdim] (lower triangle).
    Returns weights [n, symdim] and offsets [n, symdim, d]."""
    m = _c(m, np.float64)
    n, sd = m.shape
    d = {3: 2, 6: 3}[sd]
    w = np.zeros((n, sd))
    e = np.zeros((n, sd, d), dtype=np.int32)
    lib().orc_selling(d, n, _p(m), _p(w), _p(e), int(use_float))
    return w, e


def hfm_update(w, v, rhs, use_float=False, flow=False):
    """Upwind update of independent cells: w, v [n, nact], rhs [n]."""
    w, v = _c(w, np.float64), _c(v, np.float64)
    rhs = _c(np.broadcast_to(np.asarray(rhs, dtype=np.float64), (w.shape[0],)), np.float64)
    n, nact = w.shape
    u = np.zeros(n)
    fw = np.zeros((n, nact)) if flow else None
    lib().orc_hfm_update(nact, n, _p(w), _p(v), _p(rhs), _p(u), _p(fw), int(use_float))
    return (u, fw) if flow else u


# ---------------------------------------------------------------- problems
class Problem:
    """A discretised eikonal problem in the oracle's cell-major layout."""

    def __init__(self, model, shape, geom, rhs, seeds, iso_w, wall_dist=None, periodic=None):
        self.model = model if isinstance(model, int) else MODEL_IDS[model]
        self.shape = tuple(int(s) for s in shape)
        self.d = len(self.shape)
        self.N = int(np.prod(self.shape))
        self.geom = None if geom is None else _c(np.asarray(geom).reshape(-1, self.N), np.float64)
        self.rhs = _c(np.asarray(rhs).reshape(self.N), np.float64)
        self.seeds = _c(np.asarray(seeds).reshape(self.N), np.uint8)
        self.iso_w = _c(list(iso_w) + [0.] * (3 - len(iso_w)), np.float64)
        self.wall_dist = None if wall_dist is None else _c(np.asarray(wall_dist).reshape(self.N), np.uint8)
        self.periodic = _c(list(periodic or [0] * self.d) + [0] * (3 - self.d), np.int32)
        self._shape = _c(list(self.shape) + [1] * (3 - self.d), np.int64)

    @property
    def nsym(self):
        if self.model in CURVATURE_NSYM:
            return CURVATURE_NSYM[self.model]
        return self.d if self.model in (ISOTROPIC, DIAGONAL) else self.d * (self.d + 1) // 2

    @property
    def nmix(self):
        return NMIX if self.model == TTI else (2 if self.model == DUBINS2 else 1)

    def cast(self, dtype):
        """Same problem with geometry and rhs rounded to dtype (the FP32
        inputs the GPU path sees)."""
        g = None if self.geom is None else self.geom.astype(dtype).astype(np.float64)
        p = Problem(self.model, self.shape, g, self.rhs.astype(dtype).astype(np.float64), self.seeds,
                    np.asarray(self.iso_w[:self.d]).astype(dtype).astype(np.float64),
                    self.wall_dist, list(self.periodic[:self.d]))
        return p

    def solve_fm(self, u0=None):
        """Exact fixed point; with u0, that of u = min(u0, scheme(u))."""
        u = np.zeros(self.N)
        if u0 is None:
            n = lib().orc_solve_fm(self.model, self.d, _p(self._shape), _p(self.periodic), _p(self.geom),
                                   _p(self.rhs), _p(self.seeds), _p(self.wall_dist), _p(self.iso_w), _p(u))
        else:
            u0 = _c(np.asarray(u0).reshape(self.N), np.float64)
            n = lib().orc_solve_fm_from(self.model, self.d, _p(self._shape), _p(self.periodic),
                                        _p(self.geom), _p(self.rhs), _p(self.seeds), _p(self.wall_dist),
                                        _p(self.iso_w), _p(u0), _p(u))
        if n < 0:
            raise ValueError("fast marching solves the first-order scheme only (use solve_jacobi)")
        return u.reshape(self.shape)

    def solve_jacobi(self, atol, rtol, nitermax=100000, use_float=False, u0=None):
        u = np.full(self.N, np.inf) if u0 is None else _c(np.asarray(u0).reshape(self.N), np.float64).copy()
        it = lib().orc_solve_jacobi(self.model, self.d, _p(self._shape), _p(self.periodic),
                                    _p(self.geom), _p(self.rhs), _p(self.seeds), _p(self.wall_dist),
                                    _p(self.iso_w), float(atol), float(rtol), int(nitermax),
                                    int(use_float), _p(u))
        return u.reshape(self.shape), int(it)

    def solve_tiles_omp(self, atol, rtol, tile=4, niter=4, nitermax=100000):
        """Tile adaptive Gauss-Seidel on all host cores (OpenMP, FP64): the
        benchmark's multi-core CPU baseline.  Returns (u, passes, tile visits)."""
        u = np.zeros(self.N)
        visits = np.zeros(1, dtype=np.int64)
        p = lib().orc_solve_tiles_omp(self.model, self.d, _p(self._shape), _p(self.periodic),
                                      _p(self.geom), _p(self.rhs), _p(self.seeds), _p(self.wall_dist),
                                      _p(self.iso_w), int(tile), int(niter), float(atol), float(rtol),
                                      int(nitermax), _p(visits), _p(u))
        if p < 0:
            raise ValueError("the tile solver runs the first-order scheme only")
        return u.reshape(self.shape), int(p), int(visits[0])

    def stencils(self, use_float=False, with_mix=False):
        """Per-cell weights [N, nmix*nsym], offsets [N, nmix*nsym, d] (and the
        min/max mix flag [N] with with_mix)."""
        ns = self.nsym * self.nmix
        w = np.zeros((self.N, ns))
        e = np.zeros((self.N, ns, self.d), dtype=np.int32)
        mix = np.zeros(self.N, dtype=np.int32)
        lib().orc_stencils(self.model, self.d, _p(self._shape), _p(self.geom), _p(self.iso_w),
                           int(use_float), _p(w), _p(e), _p(mix))
        return (w, e, mix) if with_mix else (w, e)

    def flow(self, u, use_float=False):
        """One flow pass from values u: returns dict(values, weights [N,ns],
        indices [N,ns] (-1 for zero weight), vector [d,*shape], weightsum)."""
        ns = self.nsym
        uu = _c(np.asarray(u).reshape(self.N), np.float64).copy()
        fw = np.zeros((self.N, ns))
        ni = np.zeros((self.N, ns), dtype=np.int64)
        fv = np.zeros((self.d, self.N))
        ws = np.zeros(self.N)
        lib().orc_flow(self.model, self.d, _p(self._shape), _p(self.periodic), _p(self.geom),
                       _p(self.rhs), _p(self.seeds), _p(self.wall_dist), _p(self.iso_w),
                       int(use_float), _p(uu), _p(fw), _p(ni), _p(fv), _p(ws))
        return dict(values=uu.reshape(self.shape), weights=fw, indices=ni,
                    vector=fv.reshape((self.d,) + self.shape), weightsum=ws.reshape(self.shape))

    def adjoint(self, u, flow, sens):
        """Exact solution of (D - W)^T lam = sens, sens [*shape, nrhs]."""
        sens = np.asarray(sens, dtype=np.float64)
        nrhs = sens.shape[-1]
        s = _c(np.moveaxis(sens, -1, 0).reshape(nrhs, self.N), np.float64)
        lam = np.zeros((nrhs, self.N))
        uu = _c(np.asarray(u).reshape(self.N), np.float64)
        lib().orc_adjoint(self.N, self.nsym, _p(uu), _p(_c(flow['weights'], np.float64)),
                          _p(_c(flow['indices'], np.int64)), nrhs, _p(s), _p(lam))
        return np.moveaxis(lam.reshape((nrhs,) + self.shape), 0, -1)

    def forward_ad(self, u, flow, rhs):
        """Exact solution of (D - W) du = rhs, rhs [*shape, nrhs]."""
        rhs = np.asarray(rhs, dtype=np.float64)
        nrhs = rhs.shape[-1]
        r = _c(np.moveaxis(rhs, -1, 0).reshape(nrhs, self.N), np.float64)
        du = np.zeros((nrhs, self.N))
        uu = _c(np.asarray(u).reshape(self.N), np.float64)
        lib().orc_forward_ad(self.N, self.nsym, _p(uu), _p(_c(flow['weights'], np.float64)),
                             _p(_c(flow['indices'], np.int64)), nrhs, _p(r), _p(du))
        return np.moveaxis(du.reshape((nrhs,) + self.shape), 0, -1)


def chart_paste(u, mapping):
    """ChartPaste.h:99-165 on a cell-major grid: mapping[ndim_s, *shape[-ndim_s:]]
    gives, for the last ndim_s axes, the continuous index of the point each
    cell is identified with (broadcast over the first axes).  Returns the
    values lowered to the multilinear interpolation there where smaller
    (points within 1e-4 of the last layer moved onto it, points beyond and
    NaN skipped, no wrapping; a zero-weight infinite corner gives NaN, hence no
    paste), and the mask of lowered cells.  FP64."""
    u = np.asarray(u, dtype=np.float64)
    shape = u.shape
    m = np.asarray(mapping, dtype=np.float64)
    ns = m.shape[0]
    nb = u.ndim - ns
    m = np.broadcast_to(m.reshape((ns,) + (1,) * nb + shape[nb:]), (ns,) + shape)
    ok = np.all(np.isfinite(m), axis=0)
    q = np.floor(np.where(np.isfinite(m), m, 0.)).astype(np.int64)
    r = np.where(np.isfinite(m), m, 0.) - q
    for i in range(ns):
        qmax = shape[nb + i] - 1
        hi = (q[i] == qmax) & (r[i] <= 1e-4)
        q[i] = np.where(hi, q[i] - 1, q[i])
        r[i] = np.where(hi, 1., r[i])
        lo = (q[i] == -1) & (r[i] >= 1 - 1e-4)
        q[i] = np.where(lo, q[i] + 1, q[i])
        r[i] = np.where(lo, 0., r[i])
        ok &= (q[i] >= 0) & (q[i] < qmax)
    lead = np.indices(shape)[:nb]
    val = np.zeros(shape)
    wsum = np.zeros(shape)
    with np.errstate(invalid='ignore'):
        for c in range(1 << ns):
            w = np.ones(shape)
            idx = list(lead)
            for i in range(ns):
                e = (c >> i) & 1
                w = w * (r[i] if e else 1. - r[i])
                idx.append(np.clip(q[i] + e, 0, shape[nb + i] - 1))
            val = val + w * u[tuple(idx)]
            wsum = wsum + w
        cand = np.where(ok & (wsum > 0), val / np.where(wsum > 0, wsum, 1.), np.inf)
        lower = cand < u
    return np.where(lower, cand, u), lower


def visible(wall_dist, x, e, periodic=None):
    """The oracle's wall visibility walk (Walls.h:18-50) for the segments
    [x_q, x_q + e_q]: x, e integer arrays [n, d] on the grid of wall_dist."""
    wall_dist = _c(wall_dist, np.uint8)
    shape = wall_dist.shape
    d = len(shape)
    x = _c(np.asarray(x).reshape(-1, d), np.int64)
    e = _c(np.asarray(e).reshape(-1, d), np.int32)
    out = np.zeros(len(x), dtype=np.uint8)
    per = _c(list(periodic or [0] * d) + [0] * (3 - d), np.int32)
    lib().orc_visible(d, _p(_c(list(shape) + [1] * (3 - d), np.int64)), _p(per), _p(wall_dist),
                      len(x), _p(x), _p(e), _p(out))
    return out.astype(bool)


def wall_distance(walls, bound=10, periodic=None):
    walls = _c(walls, np.uint8)
    shape = walls.shape
    d = len(shape)
    out = np.zeros(shape, dtype=np.uint8)
    sh = _c(list(shape) + [1] * (3 - d), np.int64)
    per = _c(list(periodic or [0] * d) + [0] * (3 - d), np.int32)
    lib().orc_wall_distance(d, _p(sh), _p(per), _p(walls), int(bound), _p(out))
    return out


# ---------------------------------------------------------------- preprocessing
class curvature_relax:
    """Context: relaxation eps^2 of the curvature models' self outer products
    (0.01 by default, eps = 0.1)."""

    def __init__(self, relax):
        self.relax = float(relax)

    def __enter__(self):
        lib().orc_set_curvature_relax(self.relax)
        return self

    def __exit__(self, *exc):
        lib().orc_set_curvature_relax(0.01)


class convex:
    """Context: the Convex variants of the curvature models (left turns only)."""

    def __enter__(self):
        lib().orc_set_curvature_convex(1)
        return self

    def __exit__(self, *exc):
        lib().orc_set_curvature_convex(0)


def fejer(n=5):
    """The oracle's n-point Fejer rule on [0, pi] with weight sin: rows
    (cos phi_l, sin phi_l, w_l)."""
    out = np.zeros((n, 3))
    for l in range(n):
        lib().orc_fejer(l, n, _p(out[l]))
    return out


def curvature_geometry(shape, h, xi, kappa=0., theta=None):
    """[1/xi, kappa, cos theta, sin theta] per cell in grid units, component
    major (_SetGeometry.py:83-109): 1/xi and kappa divided by h_theta / h;
    theta defaults to the angle 2 pi k / n_theta of the third axis."""
    shape = tuple(shape)
    ratio = float(h[2]) / float(h[0])
    def field(a, name):
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 0 or a.shape == shape:
            return np.broadcast_to(a, shape)
        if a.shape == shape[2:]:
            return np.broadcast_to(a.reshape((1, 1) + shape[2:]), shape)
        if a.shape == shape[:2]:
            return np.broadcast_to(a.reshape(shape[:2] + (1,)), shape)
        raise ValueError(f"{name} has shape {a.shape}")
    if theta is None:
        theta = 2 * np.pi * np.arange(shape[2]) / shape[2]
    th = field(theta, 'theta')
    return np.stack([1. / (field(xi, 'xi') * ratio), field(kappa, 'kappa') / ratio,
                     np.cos(th), np.sin(th)]).reshape(4, -1)


class order:
    """Context manager selecting the scheme order of the oracle (2 = the
    second-order correction of HFM.h:91-134; solve with solve_jacobi)."""

    def __init__(self, k, threshold=0.3):
        self.k, self.threshold = int(k), float(threshold)

    def __enter__(self):
        lib().orc_set_order(self.k, self.threshold)
        return self

    def __exit__(self, *exc):
        lib().orc_set_order(1, 0.3)


def tti_diags(l, q, use_float=False):
    """TTI_.h:76-107 per cell: l[n, 2], q[n, 3] -> (diag[n, 7, 2], mix_is_min[n])."""
    l = _c(np.asarray(l, dtype=np.float64).reshape(-1, 2), np.float64)
    q = _c(np.asarray(q, dtype=np.float64).reshape(-1, 3), np.float64)
    n = len(l)
    diag = np.zeros((n, NMIX, 2))
    mix = np.zeros(n, dtype=np.int32)
    lib().orc_tti_diags(n, _p(l), _p(q), _p(diag), _p(mix), int(use_float))
    return diag, mix.astype(bool)


def tti_geometry(metric, h, shape):
    """TTI.with_costs(h).flatten(transposed_transformation=True)
    (Metrics/Seismic/tti.py:113-125, Metrics/base.py:175-185,
    Seismic/implicit_base.py:57-63): [l0, l1, q00, q10, q11, T] with
    T = (A diag(h))^{-T} row-major, A the inverse transformation (identity
    when absent).  metric = (linear[2, ...], quadratic[2, 2, ...], A or None)."""
    linear, quadratic, A = metric
    d = len(shape)
    full = lambda x, depth: np.broadcast_to(np.asarray(x, dtype=np.float64).reshape(
        np.shape(x) + (1,) * (d + depth - np.ndim(x))), np.shape(x)[:depth] + shape)
    l = full(linear, 1)
    q = full(quadratic, 2)
    if A is None:
        A = np.eye(d)
    A = full(A, 2)
    Ah = A * h.reshape((1, d) + (1,) * d)            # A diag(h): column j scaled by h_j
    inv = np.moveaxis(np.linalg.inv(np.moveaxis(Ah, (0, 1), (-2, -1))), (-2, -1), (0, 1))
    T = np.swapaxes(inv, 0, 1)                         # transpose of the inverse
    return np.concatenate([l, flatten_lower(q), T.reshape((d * d,) + shape)], axis=0)


def flatten_lower(m):
    d = m.shape[0]
    return np.stack([m[i, j] for i in range(d) for j in range(i + 1)], axis=0)


def grid_h(sides, dims):
    """Cell-centred grid on the box `sides` with `dims` points per axis
    (SetRect with dims=..., sampleBoundary=False): returns h, origin."""
    sides = np.asarray(sides, dtype=np.float64)
    dims = np.asarray(dims, dtype=np.float64)
    h = (sides[:, 1] - sides[:, 0]) / dims
    return h, sides[:, 0]


def set_rect(sides, float_t=np.float32, sample_boundary=False, gridScale=None, gridScales=None,
             dimx=None, dims=None):
    """Box domain of agd/Eikonal/DictIn.py:404-440, in the arithmetic type
    float_t and in the reference's order of operations (so that dims and
    origin are reproduced to the bit): the grid scale h comes from the one
    argument given, then ratios = (hi - lo)/h + sb, dims = round(ratios),
    origin = lo + ((ratios - dims) - sb) h / 2.  Returns a dict of the keys
    SetRect stores."""
    box = np.asarray(sides, dtype=float_t)
    lo, hi = box[:, 0], box[:, 1]
    sb = float(sample_boundary)
    width = hi - lo
    out = {}
    if gridScale is not None:
        out['gridScale'] = np.asarray(gridScale, dtype=float_t)
        h = np.asarray([gridScale] * len(lo), dtype=float_t)
    elif gridScales is not None:
        out['gridScales'] = np.asarray(gridScales, dtype=float_t)
        h = out['gridScales']
    elif dimx is not None:
        g = width[0] / (dimx - sb)
        out['gridScale'] = np.asarray(g, dtype=float_t)
        h = np.asarray([g] * len(lo), dtype=float_t)
    elif dims is not None:
        out['gridScales'] = width / (np.asarray(dims, dtype=float_t) - sb)
        h = out['gridScales']
    else:
        raise ValueError("no grid scale")
    ratios = (hi - lo) / h + sb
    n = np.round(ratios)
    out['dims'] = n
    out['origin'] = lo + (ratios - n - sb) * h / 2
    return out


def point_from_index(origin, h, index, float_t=np.float32, to=False):
    """DictIn.py:442-453 in float_t: x = (origin + h/2) + h i, and the
    continuous index (x - (origin + h/2)) / h for to=True."""
    origin = np.asarray(origin, dtype=float_t)
    h = np.asarray(h, dtype=float_t)
    start = origin + float_t(0.5) * h
    index = np.asarray(index, dtype=float_t)
    return (index - start) / h if to else start + h * index


def centred_axis(a, b, n):
    """The n cell centres of [a, b] as numpy.linspace(a, b, n, endpoint=False)
    shifted by half its step (DictIn.py:97-107)."""
    r, dr = np.linspace(a, b, n, endpoint=False, retstep=True)
    return r + dr / 2


def tile_expand(arr, shape_i, pad):
    """Tiled layout of a cell-major array [..., *shape] -> [..., *shape_o,
    *shape_i] (the bi-level layout of FiniteDifferences.py:164-195): pad the
    grid up to whole tiles with `pad`, split every axis k into (tile index,
    position in the tile), then gather the tile axes first.  Layout checker of
    the device's agd_block_expand."""
    d = len(shape_i)
    lead, shape = arr.shape[:-d], arr.shape[-d:]
    shape_o = [-(-s // t) for s, t in zip(shape, shape_i)]
    full = np.full(lead + tuple(o * t for o, t in zip(shape_o, shape_i)), pad, dtype=arr.dtype)
    full[(Ellipsis,) + tuple(slice(0, s) for s in shape)] = arr
    split = full.reshape(lead + tuple(x for o, t in zip(shape_o, shape_i) for x in (o, t)))
    nl = len(lead)
    perm = list(range(nl)) + [nl + 2 * k for k in range(d)] + [nl + 2 * k + 1 for k in range(d)]
    return split.transpose(perm)


def tile_squeeze(arr, shape):
    """Inverse of tile_expand: [..., *shape_o, *shape_i] -> [..., *shape]."""
    d = len(shape)
    nl = arr.ndim - 2 * d
    perm = list(range(nl)) + [x for k in range(d) for x in (nl + k, nl + d + k)]
    inter = arr.transpose(perm)
    lead = arr.shape[:nl]
    full = inter.reshape(lead + tuple(arr.shape[nl + k] * arr.shape[nl + d + k] for k in range(d)))
    return full[(Ellipsis,) + tuple(slice(0, s) for s in shape)]


def seed_indices(seeds, origin, h, dtype=np.float32):
    """Continuous index (x - origin)/h - 1/2 in dtype, rounded half to even."""
    seeds = np.asarray(seeds, dtype=dtype)
    start = np.asarray(origin, dtype=dtype) + dtype(0.5) * np.asarray(h, dtype=dtype)
    cont = (seeds - start) / np.asarray(h, dtype=dtype)
    return np.round(cont).astype(np.int64)


def make_problem(model, shape, h, metric=None, drift=None, cost=1., seeds_idx=(), seed_values=None,
                 walls=None, wall_bound=10, periodic=None):
    """Cell-major FP64 problem, restating _SetGeometry.py / _SetArgs.py.

    metric: Riemann/Rander primal matrices m[d, d, *shape] (or [d, d]);
    Diagonal: costs[d, *shape]; ignored for Isotropic.  drift: Rander w[d, ...].
    """
    shape = tuple(int(s) for s in shape)
    d = len(shape)
    N = int(np.prod(shape))
    h = np.asarray(h, dtype=np.float64)
    mid = model if isinstance(model, int) else MODEL_IDS[model]
    geom = None
    iso_w = [0.] * d
    if mid == ISOTROPIC:
        iso_w = list(h ** -2)
    elif mid == DIAGONAL:
        c = np.asarray(metric, dtype=np.float64)
        if c.ndim == 1:
            c = c.reshape((d,) + (1,) * d)
        c = np.broadcast_to(c, (d,) + shape)
        geom = (c * h.reshape((d,) + (1,) * d)) ** -2
    elif mid == TTI:
        geom = tti_geometry(metric, h, shape)
    else:
        m = np.asarray(metric, dtype=np.float64)
        if m.ndim == 2:
            m = np.broadcast_to(m.reshape(d, d, *(1,) * d), (d, d) + shape)
        mh = m * (h[:, None] * h[None, :]).reshape((d, d) + (1,) * d)
        inv = np.moveaxis(np.linalg.inv(np.moveaxis(mh, (0, 1), (-2, -1))), (-2, -1), (0, 1))
        geom = flatten_lower(inv)
        if mid == RANDER:
            w = np.asarray(drift, dtype=np.float64)
            if w.ndim == 1:
                w = np.broadcast_to(w.reshape(d, *(1,) * d), (d,) + shape)
            geom = np.concatenate([geom, w * h.reshape((d,) + (1,) * d)], axis=0)
    rhs = np.broadcast_to(np.asarray(cost, dtype=np.float64), shape).copy()
    seeds = np.zeros(shape, dtype=np.uint8)
    idx = np.asarray(seeds_idx, dtype=np.int64).reshape(-1, d)
    vals = np.zeros(len(idx)) if seed_values is None else np.asarray(seed_values, dtype=np.float64)
    for i, s in enumerate(idx):
        rhs[tuple(s)] = vals[i]
        seeds[tuple(s)] = 1
    wd = None
    if walls is not None:
        walls = np.asarray(walls, dtype=bool)
        rhs[walls] = np.inf
        seeds[walls] = 1
        wd = wall_distance(walls, wall_bound, periodic)
    return Problem(mid, shape, None if geom is None else geom.reshape(-1, N), rhs, seeds, iso_w,
                   wd, periodic)


# ---------------------------------------------------------------- source factorization
class factor:
    """Context manager: source factorization (EuclideanFactor.h:8-50,
    Constants.h:50-61, TTI_.h:139-212) for the following solves and flows:
    radius in grid units around origin (grid units), metric m at the origin in
    grid units (squared costs for Isotropic/Diagonal, lower triangle for
    Riemann/Rander, [l0, l1, q00, q10, q11, A row-major] for TTI; see
    factor_metric)."""

    def __init__(self, radius, origin, m):
        self.r2 = float(radius) ** 2
        self.origin = _c(list(np.asarray(origin, dtype=np.float64).reshape(-1)) + [0.] * 3, np.float64)[:3].copy()
        self.m = _c(list(np.asarray(m, dtype=np.float64).reshape(-1)) + [0.] * 14, np.float64)[:14].copy()

    def __enter__(self):
        lib().orc_set_factor(self.r2, _p(self.origin), _p(self.m))
        return self

    def __exit__(self, *exc):
        lib().orc_set_factor(0., None, None)


def _interp_corners(pts, shape, periodic):
    """Multilinear (order 1) interpolation stencil at points pts[k, d] (grid
    units): corner cells [2^d, k, d] and weights [2^d, k].  Non-periodic axes
    clamp the point onto [0, n-1]."""
    pts = np.asarray(pts, dtype=np.float64)
    k, d = pts.shape
    lo = np.zeros((k, d), dtype=np.int64)
    t = np.zeros((k, d))
    for c in range(d):
        n = shape[c]
        x = pts[:, c]
        if periodic[c]:
            b = np.floor(x)
        else:
            x = np.clip(x, 0, n - 1)
            b = np.clip(np.floor(x), 0, max(n - 2, 0))
        lo[:, c] = b.astype(np.int64)
        t[:, c] = x - b
    cells, weights = [], []
    for bits in range(2 ** d):
        off = np.array([(bits >> (d - 1 - c)) & 1 for c in range(d)])
        idx = lo + off
        w = np.prod(np.where(off == 1, t, 1 - t), axis=1)
        for c in range(d):
            idx[:, c] = idx[:, c] % shape[c] if periodic[c] else np.minimum(idx[:, c], shape[c] - 1)
        cells.append(idx)
        weights.append(w)
    return np.stack(cells), np.stack(weights)


def cost_metric_at(model, pts, h, shape, metric=None, drift=None, cost=1., periodic=None):
    """The cost metric of _Interface.py:202-211 in grid units at points
    pts[k, d] (grid units): metric.with_costs(h).with_cost(cost), i.e.
    Isotropic c_i = h_i cost, Diagonal c_i = costs_i h_i cost, Riemann
    m_ij h_i h_j cost^2, Rander (m_ij h_i h_j cost^2, w_i h_i cost),
    multilinearly interpolated (Interpolation.py order 1).  Returns
    ('diag', c[k, d]) or ('sym', m[k, d, d], w[k, d] or None)."""
    mid = model if isinstance(model, int) else MODEL_IDS[model]
    shape = tuple(int(s) for s in shape)
    d = len(shape)
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, d)
    h = np.asarray(h, dtype=np.float64)
    periodic = list(periodic or [0] * d)
    cells, wts = _interp_corners(pts, shape, periodic)

    def field(a, core):
        a = np.asarray(a, dtype=np.float64)
        if a.shape == core:
            a = a.reshape(core + (1,) * d)
        return np.broadcast_to(a, core + shape)

    def at(f, core):
        # f[core..., *shape] at every corner -> [2^d, k, *core]
        sel = f[(Ellipsis,) + tuple(np.moveaxis(cells, -1, 0))]
        return np.moveaxis(sel, tuple(range(len(core))), tuple(range(2, 2 + len(core))))

    c = at(field(cost, ()), ())
    if mid == TTI:
        # metric = (linear (2, ...), quadratic (2, 2, ...), A (d, d, ...) or None):
        # with_costs(h) then with_cost(cost) scale A's columns by h and A by the
        # cost (Seismic/implicit_base.py:48-56); each field is interpolated.
        lin, quad, A = metric
        li = np.einsum('ck,ckd->kd', wts, at(field(lin, (2,)), (2,)))
        qf = np.asarray(quad, dtype=np.float64)
        q3 = np.stack([qf[0, 0], qf[1, 0], qf[1, 1]])
        qi = np.einsum('ck,ckd->kd', wts, at(field(q3, (3,)), (3,)))
        A = np.eye(d) if A is None else A
        Ah = at(field(A, (d, d)), (d, d)) * h[None, None, None, :] * c[..., None, None]
        return 'tti', li, qi, np.einsum('ck,ckij->kij', wts, Ah)
    if mid in (ISOTROPIC, DIAGONAL):
        costs = np.ones((d,)) if mid == ISOTROPIC else metric
        cc = at(field(costs, (d,)), (d,)) * h * c[..., None]
        return 'diag', np.einsum('ck,ckd->kd', wts, cc)
    if mid not in (RIEMANN, RANDER):
        raise NotImplementedError("cost metric of the TTI model")
    m = at(field(metric, (d, d)), (d, d)) * np.outer(h, h) * (c ** 2)[..., None, None]
    mi = np.einsum('ck,ckij->kij', wts, m)
    wi = None
    if mid == RANDER:
        w = at(field(drift, (d,)), (d,)) * h * c[..., None]
        wi = np.einsum('ck,ckd->kd', wts, w)
    return 'sym', mi, wi


def tti_primal_norm(l, q3, A, v, nsample=4001):
    """The TTI norm N(v) = N0(A v), N0 the support function of the dual ball
    l0 P0 + l1 P1 + (q00 P0^2 + 2 q10 P0 P1 + q11 P1^2) / 2 <= 1 with
    P0 = X^2 [+ Y^2], P1 = Z^2 (Seismic/tti.py:64-69), by brute force: the
    ball's boundary in the quarter plane (rho, z) = t (cos th, sin th), t^2
    the smallest positive root of the quadratic along the ray, maximising
    rho |w_lateral| + z |w_last| over th on a grid and then by golden-section
    refinement of the best bracket.  l[k, 2], q3[k, 3], A[k, d, d], v[k, d]."""
    l = np.atleast_2d(np.asarray(l, dtype=np.float64))
    q3 = np.atleast_2d(np.asarray(q3, dtype=np.float64))
    A = np.asarray(A, dtype=np.float64)
    v = np.atleast_2d(np.asarray(v, dtype=np.float64))
    d = v.shape[1]
    A = np.broadcast_to(A.reshape((-1, d, d)), (len(v), d, d))
    w = np.einsum('kij,kj->ki', A, v)
    r = np.sqrt(np.sum(w[:, :d - 1] ** 2, axis=1))
    z = np.abs(w[:, d - 1])
    l = np.broadcast_to(l, (len(v), 2))
    q3 = np.broadcast_to(q3, (len(v), 3))

    def value(th):  # th[k, m]
        c2, s2 = np.cos(th) ** 2, np.sin(th) ** 2
        L = l[:, :1] * c2 + l[:, 1:] * s2
        Q = q3[:, :1] * c2 * c2 + 2. * q3[:, 1:2] * c2 * s2 + q3[:, 2:] * s2 * s2
        tau = 2. / (L + np.sqrt(L * L + 2. * Q))  # root of Q tau^2 / 2 + L tau - 1
        t = np.sqrt(tau)
        return t * (np.cos(th) * r[:, None] + np.sin(th) * z[:, None])

    th = np.linspace(0., np.pi / 2, nsample)[None, :].repeat(len(v), 0)
    vals = value(th)
    j = np.argmax(vals, axis=1)
    step = th[0, 1] - th[0, 0]
    lo = np.clip(th[np.arange(len(v)), j] - step, 0., np.pi / 2)
    hi = np.clip(th[np.arange(len(v)), j] + step, 0., np.pi / 2)
    g = (np.sqrt(5.) - 1.) / 2.
    for _ in range(80):
        a = hi - g * (hi - lo)
        b = lo + g * (hi - lo)
        fa, fb = value(a[:, None])[:, 0], value(b[:, None])[:, 0]
        keep_lo = fa > fb
        hi = np.where(keep_lo, b, hi)
        lo = np.where(keep_lo, lo, a)
    best = value((0.5 * (lo + hi))[:, None])[:, 0]
    return np.maximum(best, vals.max(axis=1))


def tti_factor(d, fm, x, e=None, f32=False, order2=False):
    """The TTI factorization of the oracle at points x[n, d] (relative to the
    origin) for the metric fm = [l0, l1, q00, q10, q11, A row-major]: returns
    (norm[n], grad[n, d]) and, with offsets e[n, d], also fact[n, 2] (shifts
    of x - e, x + e) and, with order2, fact2[n, 2] (x - 2e, x + 2e)."""
    x = _c(np.atleast_2d(x), np.float64)
    n = len(x)
    fmv = _c(list(np.asarray(fm, dtype=np.float64).reshape(-1)) + [0.] * 14, np.float64)[:14].copy()
    norm = np.zeros(n)
    grad = np.zeros((n, d))
    ee = fact = fact2 = None
    if e is not None:
        ee = np.zeros((n, 3), dtype=np.int32)
        ee[:, :d] = np.asarray(e).reshape(n, d)
        fact = np.zeros((n, 2))
        fact2 = np.zeros((n, 2)) if order2 else None
    lib().orc_tti_factor(d, int(f32), _p(fmv), n, _p(x), _p(ee), _p(norm), _p(grad), _p(fact), _p(fact2))
    if e is None:
        return norm, grad
    return (norm, grad, fact, fact2) if order2 else (norm, grad, fact)


def metric_norm(cm, diff):
    """Norms of diff[k, d] for the cost metric cm (one point, broadcast, or
    one per row)."""
    diff = np.asarray(diff, dtype=np.float64)
    if cm[0] == 'tti':
        return tti_primal_norm(cm[1], cm[2], cm[3], diff)
    if cm[0] == 'diag':
        return np.sqrt(np.sum((cm[1] * diff) ** 2, axis=-1))
    m, w = cm[1], cm[2]
    n = np.sqrt(np.einsum('ki,kij,kj->k', diff, np.broadcast_to(m, (len(diff),) + m.shape[-2:]), diff))
    if w is not None:
        n = n + np.sum(np.broadcast_to(w, diff.shape) * diff, axis=-1)
    return n


def factor_metric(cm):
    """The factorization metric of _Kernel.py:301-308 from the cost metric at
    the seed: squared costs (diagonal models), [l, q, A] (TTI), else the
    lower triangle of m (the drift of Rander is left out)."""
    if cm[0] == 'diag':
        return (cm[1].reshape(-1) ** 2)
    if cm[0] == 'tti':
        A = cm[3].reshape(cm[3].shape[-2:])
        return np.concatenate([cm[1].reshape(-1), cm[2].reshape(-1), A.reshape(-1)])
    m = cm[1].reshape(cm[1].shape[-2:])
    return flatten_lower(m)


def spread_seed(seed, radius, shape, cm_at, value=0., periodic=None):
    """_SetArgs.py:46-75: the single seed (grid units) is replaced by the
    cells at Euclidean distance < radius, inside the domain on non-periodic
    axes, with values value + (|x - seed|_{M(seed)} + |x - seed|_{M(x)}) / 2
    for the cost metric M = cm_at(points).  Returns (cells[k, d], values[k])."""
    seed = np.asarray(seed, dtype=np.float64).reshape(-1)
    d = len(seed)
    periodic = np.asarray(list(periodic or [0] * d), dtype=bool)
    axes = [np.arange(int(np.floor(c - radius)), int(np.ceil(c + radius)) + 1) for c in seed]
    neigh = np.stack(np.meshgrid(*axes, indexing='ij'), -1).reshape(-1, d)
    neigh = neigh[np.linalg.norm(neigh - seed, axis=1) < radius]
    shp = np.asarray(shape)
    nper = ~periodic
    inside = np.all((neigh[:, nper] >= -0.5) & (neigh[:, nper] < shp[nper] - 0.5), axis=1)
    neigh = neigh[inside]
    diff = neigh - seed
    vals = value + 0.5 * (metric_norm(cm_at(seed[None]), diff) + metric_norm(cm_at(neigh), diff))
    cells = np.where(periodic, neigh % shp, neigh)
    return cells.astype(np.int64), vals
