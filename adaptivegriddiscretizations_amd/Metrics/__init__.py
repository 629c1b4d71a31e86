"""Norm classes accepted by the eikonal solver's 'metric' / 'dualMetric' keys.

Host-side containers for the metric fields, with the methods of the reference
classes that the GPU eikonal path relies on (agd/Metrics/base.py,
isotropic.py, diagonal.py, riemann.py, rander.py).  Arrays may be numpy
arrays or torch tensors; the algebra helpers (dual, with_costs, flatten, ...)
work on numpy arrays.  The solver itself applies the grid scale and inverts
the metric on the GPU (agd_prepare_geometry); these methods exist so that
user code written against the reference keeps working.
"""
import numpy as np

__all__ = ["Base", "Isotropic", "Diagonal", "Riemann", "Rander", "Seismic"]


def _np(a):
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def flatten_symmetric_matrix(m):
    """Lower triangular entries m[i, :i+1], concatenated (misc.py:8-15)."""
    d = m.shape[0]
    assert d == m.shape[1]
    return np.concatenate([m[i, :(i + 1)] for i in range(d)], axis=0)


def expand_symmetric_matrix(arr, d=None, extra_length=False):
    """Inverse of flatten_symmetric_matrix (misc.py:17-28)."""
    if d is None:
        d = 0
        while (d * (d + 1)) // 2 < len(arr):
            d += 1
    assert extra_length or len(arr) == (d * (d + 1)) // 2

    def index(i, j):
        i, j = max(i, j), min(i, j)
        return (i * (i + 1)) // 2 + j
    return np.asarray([[arr[index(i, j)] for i in range(d)] for j in range(d)])


def _inverse(m):
    """Inverse of a field of matrices of shape (d, d, *shape)."""
    m = np.asarray(m)
    return np.moveaxis(np.linalg.inv(np.moveaxis(m, (0, 1), (-2, -1))), (-2, -1), (0, 1))


def _outer_self(v):
    return v[:, None] * v[None, :]


class Base:
    """Common interface (base.py:175-210)."""

    def with_costs(self, costs):
        raise NotImplementedError

    def with_speeds(self, speeds):
        return self.with_costs(1. / np.asarray(speeds))

    def with_cost(self, cost):
        cost = np.asarray(cost)
        return self.with_costs(np.broadcast_to(cost, (self.vdim,) + cost.shape))

    def with_speed(self, speed):
        return self.with_cost(1. / np.asarray(speed))

    @classmethod
    def from_HFM(cls, arr):
        return cls.expand(arr)


class Isotropic(Base):
    """F(x) = cost |x| (isotropic.py)."""

    def __init__(self, cost, vdim=None):
        self.cost = cost
        self._vdim = vdim

    @classmethod
    def from_speed(cls, speed, vdim=None):
        return cls(1. / np.asarray(speed), vdim)

    @property
    def vdim(self):
        if self._vdim is None:
            raise ValueError("Isotropic metric: unspecified dimension")
        return self._vdim

    @property
    def shape(self):
        return np.shape(self.cost)

    def dual(self):
        return Isotropic(1. / np.asarray(self.cost), self._vdim)

    def norm(self, v):
        return np.asarray(self.cost) * np.linalg.norm(np.asarray(v), axis=0)

    def flatten(self):
        return np.asarray(self.cost)

    def with_costs(self, costs):
        costs = np.asarray(costs)
        if np.all(costs == costs[0]):
            return Isotropic(np.asarray(self.cost) * costs[0], self._vdim)
        return Diagonal(np.asarray(self.cost) * costs)

    @classmethod
    def expand(cls, arr):
        return cls(arr)


class Diagonal(Base):
    """F(x) = |costs * x| (diagonal.py)."""

    def __init__(self, costs):
        self.costs = costs

    @classmethod
    def from_speed(cls, speeds):
        return cls(1. / np.asarray(speeds))

    @property
    def vdim(self):
        return len(self.costs)

    @property
    def shape(self):
        return tuple(np.shape(self.costs)[1:])

    def dual(self):
        return Diagonal.from_speed(self.costs)

    def with_costs(self, costs):
        costs = np.asarray(costs)
        c = np.asarray(self.costs)
        costs = costs.reshape(costs.shape + (1,) * (c.ndim - costs.ndim))
        return Diagonal(c * costs)

    def norm(self, v):
        c = np.asarray(self.costs)
        v = np.asarray(v)
        return np.linalg.norm(c.reshape(c.shape + (1,) * (v.ndim - c.ndim)) * v, axis=0)

    def flatten(self):
        return self.costs

    @classmethod
    def expand(cls, arr):
        return cls(arr)


class Riemann(Base):
    """F(x) = sqrt(<x, m x>), m of shape (d, d, *shape) (riemann.py)."""

    def __init__(self, m):
        self.m = m

    @property
    def vdim(self):
        return len(self.m)

    @property
    def shape(self):
        return tuple(np.shape(self.m)[2:])

    def dual(self):
        return Riemann(_inverse(_np(self.m)))

    def with_costs(self, costs):
        m = _np(self.m)
        costs = np.asarray(costs)
        costs = costs.reshape(costs.shape + (1,) * (m.ndim - 1 - costs.ndim))
        return Riemann(m * _outer_self(costs))

    def norm(self, v):
        m = _np(self.m)
        v = np.asarray(v)
        m = m.reshape(m.shape + (1,) * (v.ndim + 1 - m.ndim))
        return np.sqrt(np.einsum("i...,ij...,j...->...", v, m, v))

    def flatten(self):
        return flatten_symmetric_matrix(_np(self.m))

    @classmethod
    def expand(cls, arr):
        return cls(expand_symmetric_matrix(np.asarray(arr)))

    @classmethod
    def needle(cls, u, cost_parallel, cost_orthogonal):
        """Eigenvector u with eigenvalue cost_parallel**2, orthogonal
        eigenvalue cost_orthogonal**2 (riemann.py:80-105)."""
        u = np.asarray(u, dtype=float)
        nu = np.linalg.norm(u, axis=0)
        u = np.where(nu > 0, u / np.where(nu > 0, nu, 1), u)
        cp_, co_ = np.asarray(cost_parallel), np.asarray(cost_orthogonal)
        d = len(u)
        ident = np.eye(d).reshape((d, d) + (1,) * (u.ndim - 1))
        return cls((cp_ ** 2 - co_ ** 2) * _outer_self(u) + co_ ** 2 * ident)

    @classmethod
    def from_diagonal(cls, diag):
        diag = np.asarray(diag)
        d = len(diag)
        z = np.zeros_like(diag[0])
        return cls(np.asarray([[diag[i] if i == j else z for i in range(d)] for j in range(d)]))

    @classmethod
    def from_cast(cls, metric):
        if isinstance(metric, cls):
            return metric
        if isinstance(metric, Diagonal):
            return cls.from_diagonal(np.asarray(metric.costs) ** 2)
        if isinstance(metric, Isotropic):
            d = metric.vdim
            c2 = np.asarray(metric.cost) ** 2
            return cls.from_diagonal(np.broadcast_to(c2, (d,) + np.shape(c2)))
        raise ValueError(f"Cannot cast {type(metric)} to Riemann")


class Rander(Base):
    """F(x) = sqrt(<x, m x>) + <w, x> (rander.py)."""

    def __init__(self, m, w):
        self.m, self.w = m, w

    @property
    def vdim(self):
        return len(self.m)

    @property
    def shape(self):
        return tuple(np.shape(self.m)[2:])

    def dual(self):
        m, w = _np(self.m), _np(self.w)
        m, w = np.broadcast_arrays(m, w[None] if w.ndim == m.ndim - 1 else w)
        w = w[0]
        s = _inverse(m - _outer_self(w))
        omega = np.einsum("ij...,j...->i...", s, w)
        return Rander((1 + np.einsum("i...,i...->...", w, omega)) * s, -omega)

    def with_costs(self, costs):
        m, w = _np(self.m), _np(self.w)
        costs = np.asarray(costs)
        c1 = costs.reshape(costs.shape + (1,) * (w.ndim - costs.ndim))
        c2 = costs.reshape(costs.shape + (1,) * (m.ndim - 1 - costs.ndim))
        return Rander(m * _outer_self(c2), w * c1)

    def norm(self, v):
        v = np.asarray(v)
        return Riemann(self.m).norm(v) + np.einsum("i...,i...->...", _np(self.w), v)

    def flatten(self, inverse_m=False):
        m = _np(self.m)
        if inverse_m:
            m = _inverse(m)
        return np.concatenate((flatten_symmetric_matrix(m), _np(self.w)), axis=0)

    @classmethod
    def expand(cls, arr):
        arr = np.asarray(arr)
        m = expand_symmetric_matrix(arr, extra_length=True)
        d = len(m)
        return cls(m, arr[d * (d + 1) // 2:])

    @classmethod
    def from_Zermelo(cls, metric, drift):
        """Zermelo navigation: unit speed for `metric`, plus `drift`
        (rander.py:81-97)."""
        return cls(_inverse(np.asarray(metric)), np.asarray(drift)).dual()

    @classmethod
    def from_cast(cls, metric):
        if isinstance(metric, cls):
            return metric
        r = Riemann.from_cast(metric)
        return cls(r.m, np.zeros(r.vdim))


from . import Seismic  # noqa: E402,F401
