"""Synthetic eikonal workloads used by the benchmark and the tests.

Each generator returns a Workload: the model name, the box and grid, the
float32 input fields and the seed points, plus `dict_in()` building the
matching Eikonal.dictIn.  Random fields come from numpy's PCG64 with the
seed given, so a workload is reproducible on any host.

  isotropic2_bumps   2D isotropic, cost 1 + 0.5 tanh(sum of 32 Gaussian
                     bumps), seed at the origin            (512^2 headline)
  riemann2_aniso     2D Riemannian, M = R(theta)^T diag(1, a^2) R(theta),
                     theta and log a smooth random fields   (2000^2 headline)
  riemann3_aniso     3D Riemannian, M = Q diag(l1,l2,l3) Q^T, Q from a smooth
                     quaternion field, log l smooth in [0, ln 5]  (200^3)
  curvature2         a curvature-penalised model (ReedsShepp2, ...) on
                     [-1,1]^2 x [0, 2 pi): cost 1 + 0.5 tanh(bumps) in the
                     plane, constant xi and kappa, seed at (0, 0, 0)
"""
from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter


@dataclass
class Workload:
    name: str
    model: str
    sides: list
    dims: tuple
    seeds: np.ndarray                 # physical seed points (n, d)
    metric: np.ndarray = None         # (d, d, *dims) float32 primal metric
    cost: np.ndarray = None           # (*dims) float32 or None (=1)
    receivers: np.ndarray = None      # (n, d) integer cells
    extra: dict = field(default_factory=dict)

    @property
    def ndim(self):
        return len(self.dims)

    @property
    def h(self):
        s = np.asarray(self.sides, dtype=np.float64)
        h = (s[:, 1] - s[:, 0]) / np.asarray(self.dims[:len(s)], dtype=np.float64)
        if len(s) < len(self.dims):      # angle axis of a curvature-penalised model
            h = np.concatenate([h, [2 * np.pi / self.dims[-1]]])
        return h

    @property
    def ncells(self):
        return int(np.prod(self.dims))

    def dict_in(self, **extra):
        from .Eikonal import dictIn
        from . import Metrics
        d = {'model': self.model, 'seeds': self.seeds, 'verbosity': 0}
        if self.metric is not None:
            d['metric'] = Metrics.Riemann(self.metric)
        if 'tti' in self.extra:
            d['metric'] = self.extra['tti']
        if self.cost is not None:
            d['cost'] = self.cost
        for key in ('xi', 'kappa'):
            if key in self.extra:
                d[key] = self.extra[key]
        d.update(extra)
        hfmIn = dictIn(d)
        if len(self.sides) < len(self.dims):
            hfmIn.SetRect(self.sides, dimx=self.dims[0])
            hfmIn.nTheta = self.dims[-1]
        else:
            hfmIn.SetRect(self.sides, dims=list(self.dims))
        return hfmIn


def _smooth_uniform(rng, shape, sigma, lo, hi):
    """Gaussian-filtered uniform noise, min-max mapped to [lo, hi)."""
    f = gaussian_filter(rng.uniform(size=shape), sigma=sigma, mode='wrap')
    f = (f - f.min()) / max(f.max() - f.min(), 1e-30)
    return lo + (hi - lo) * f * (1 - 1e-7)


def isotropic2_bumps(n=512, seed=0, nbumps=32):
    """2D isotropic eikonal on [-1,1]^2, cost c = 1 + 0.5 tanh(g), g a sum of
    Gaussian bumps (centres uniform, widths 0.05-0.2, amplitudes N(0,1)),
    single seed at the origin."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-1, 1, (nbumps, 2))
    widths = rng.uniform(0.05, 0.2, nbumps)
    amps = rng.normal(size=nbumps)
    x = -1 + (np.arange(n) + 0.5) * (2. / n)
    X = np.stack(np.meshgrid(x, x, indexing='ij'))
    g = np.zeros((n, n))
    for c, s, a in zip(centers, widths, amps):
        g += a * np.exp(-((X[0] - c[0]) ** 2 + (X[1] - c[1]) ** 2) / (2 * s ** 2))
    cost = (1 + 0.5 * np.tanh(g)).astype(np.float32)
    return Workload('isotropic2_bumps', 'Isotropic2', [[-1, 1], [-1, 1]], (n, n),
                    np.zeros((1, 2), dtype=np.float32), cost=cost)


def isotropic2_batch(ninst=256, seed=3, nmin=64, nmax=1024, nbumps=32, device=None):
    """Batch of 2D isotropic problems on [-1,1]^2 with independent sides
    uniform in {nmin..nmax} per axis (so h differs per axis and problem),
    costs c = 1 + 0.5 tanh(g) as isotropic2_bumps with per-problem bumps,
    and 1..8 seeds at random points.  Costs are built with torch on
    `device` (the grids total ~76 M cells for 256 problems)."""
    import torch
    rng = np.random.default_rng(seed)
    dev = torch.device(device) if device is not None else torch.device('cpu')
    out = []
    for i in range(ninst):
        n0, n1 = (int(v) for v in rng.integers(nmin, nmax + 1, 2))
        centers = torch.tensor(rng.uniform(-1, 1, (nbumps, 2)), device=dev)
        widths = torch.tensor(rng.uniform(0.05, 0.2, nbumps), device=dev)
        amps = torch.tensor(rng.normal(size=nbumps), device=dev)
        x0 = -1 + (torch.arange(n0, device=dev, dtype=torch.float64) + 0.5) * (2. / n0)
        x1 = -1 + (torch.arange(n1, device=dev, dtype=torch.float64) + 0.5) * (2. / n1)
        d0 = (x0[None, :] - centers[:, 0:1]) ** 2          # (nbumps, n0)
        d1 = (x1[None, :] - centers[:, 1:2]) ** 2          # (nbumps, n1)
        w = 1. / (2 * widths ** 2)
        g = torch.einsum('b,bi,bj->ij', amps, torch.exp(-d0 * w[:, None]), torch.exp(-d1 * w[:, None]))
        cost = (1 + 0.5 * torch.tanh(g)).to(torch.float32)
        nseeds = int(rng.integers(1, 9))
        seeds = rng.uniform(-1, 1, (nseeds, 2)).astype(np.float32)
        out.append(Workload(f'isotropic2_batch[{i}]', 'Isotropic2', [[-1, 1], [-1, 1]], (n0, n1),
                            seeds, cost=cost))
    return out


def riemann2_aniso(n=2000, seed=1, nseeds=4, sigma=20., amax=10.):
    """2D Riemannian eikonal on [-1,1]^2: M = R(theta)^T diag(1, a^2) R(theta),
    theta smooth in [0, pi), a log-uniform in [1, amax] and smoothed (sigma in
    cells); nseeds seeds at random cell centres."""
    rng = np.random.default_rng(seed)
    theta = _smooth_uniform(rng, (n, n), sigma, 0., np.pi)
    loga = _smooth_uniform(rng, (n, n), sigma, 0., np.log(amax))
    a2 = np.exp(2 * loga)
    c, s = np.cos(theta), np.sin(theta)
    # R = [[c, -s], [s, c]];  M = R^T diag(1, a^2) R
    m00 = c * c + a2 * s * s
    m01 = -c * s + a2 * s * c
    m11 = s * s + a2 * c * c
    metric = np.array([[m00, m01], [m01, m11]], dtype=np.float32)
    h = 2. / n
    cells = rng.integers(0, n, (nseeds, 2))
    seeds = (-1 + (cells + 0.5) * h).astype(np.float32)
    return Workload('riemann2_aniso', 'Riemann2', [[-1, 1], [-1, 1]], (n, n), seeds,
                    metric=metric, extra={'seed_cells': cells})


def tti_tilted(n=256, d=2, seed=5, sigma=None, max_tilt=0.6, hexagonal=(15., 4., 10., 3.)):
    """TTI medium on [-1,1]^d: a hexagonal material (c11, c13, c33, c44)
    tilted by a smooth rotation field (angle up to max_tilt about a smooth
    axis in 3D), one seed at the centre (TTI.from_hexagonal, tti.py:141)."""
    from .Metrics.Seismic import TTI
    rng = np.random.default_rng(seed)
    sigma = n / 16. if sigma is None else sigma
    shape = (n,) * d
    c11, c13, c33, c44 = hexagonal
    base = TTI.from_hexagonal(c11, None, c13, c33, c44, vdim=d)
    th = _smooth_uniform(rng, shape, sigma, -max_tilt, max_tilt)
    if d == 2:
        c, s = np.cos(th), np.sin(th)
        A = np.array([[c, -s], [s, c]])
    else:
        phi = _smooth_uniform(rng, shape, sigma, 0., 2 * np.pi)
        axis = np.array([np.cos(phi), np.sin(phi), np.zeros_like(phi)])   # horizontal tilt axis
        q = np.concatenate([np.cos(th / 2)[None], np.sin(th / 2)[None] * axis])
        A = rotation_from_quaternion(q)
    metric = TTI(base.linear, base.quadratic, vdim=d, inverse_transformation=A.astype(np.float32))
    seeds = np.zeros((1, d), dtype=np.float32)
    return Workload(f'tti{d}_tilted', f'TTI{d}', [[-1, 1]] * d, shape, seeds, metric=None,
                    extra={'tti': metric})


def curvature2(model='ReedsShepp2', n=64, ntheta=32, xi=0.3, kappa=0., seed=7, nbumps=16):
    """Curvature-penalised model on [-1,1]^2 x [0, 2 pi) with n x n x ntheta
    cells: a planar cost 1 + 0.5 tanh(sum of nbumps Gaussian bumps), constant
    xi and kappa, one seed at the origin heading along +x."""
    rng = np.random.default_rng(seed)
    x = -1 + (np.arange(n) + 0.5) * (2. / n)
    X, Y = np.meshgrid(x, x, indexing='ij')
    acc = np.zeros((n, n))
    for _ in range(nbumps):
        c = rng.uniform(-1, 1, size=2)
        r = rng.uniform(0.05, 0.3)
        a = rng.uniform(-2, 2)
        acc += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * r * r))
    cost = (1. + 0.5 * np.tanh(acc)).astype(np.float32)
    seeds = np.zeros((1, 3), dtype=np.float32)
    return Workload(f'{model.lower()}_{n}x{n}x{ntheta}', model, [[-1, 1], [-1, 1]], (n, n, ntheta), seeds,
                    cost=cost, extra={'xi': xi, 'kappa': kappa})


def rotation_from_quaternion(q):
    """Rotation matrices of unit quaternions q[4, ...] (r, i, j, k)."""
    r, i, j, k = q
    return np.array([
        [1 - 2 * (j * j + k * k), 2 * (i * j - k * r), 2 * (i * k + j * r)],
        [2 * (i * j + k * r), 1 - 2 * (i * i + k * k), 2 * (j * k - i * r)],
        [2 * (i * k - j * r), 2 * (j * k + i * r), 1 - 2 * (i * i + j * j)]])


def riemann3_aniso(n=200, seed=2, sigma=10., lmax=5., nreceivers=64):
    """3D Riemannian eikonal on [-1,1]^3: M = Q diag(l) Q^T, Q from a
    Gaussian-smoothed quaternion field (4 N(0,1) channels, normalised), log l
    uniform in [0, ln lmax] and smoothed; one seed at the centre cell;
    nreceivers receivers at random boundary cells."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    q = np.stack([gaussian_filter(rng.normal(size=shape), sigma=sigma, mode='wrap') for _ in range(4)])
    q /= np.linalg.norm(q, axis=0)
    Q = rotation_from_quaternion(q)
    lam = np.stack([np.exp(_smooth_uniform(rng, shape, sigma, 0., np.log(lmax))) for _ in range(3)])
    metric = np.einsum('ik...,k...,jk...->ij...', Q, lam, Q).astype(np.float32)
    h = 2. / n
    centre = np.array([[n // 2] * 3])
    seeds = (-1 + (centre + 0.5) * h).astype(np.float32)
    rec = rng.integers(0, n, (nreceivers, 3))
    axis = rng.integers(0, 3, nreceivers)
    side = rng.integers(0, 2, nreceivers)
    rec[np.arange(nreceivers), axis] = side * (n - 1)
    return Workload('riemann3_aniso', 'Riemann3', [[-1, 1]] * 3, shape, seeds, metric=metric,
                    receivers=rec, extra={'seed_cells': centre})
