"""dictIn / dictOut: the input and output containers of the eikonal solver.

Drop-in counterpart of agd/Eikonal/DictIn.py for the GPU path: the same keys
and key aliases, the same grid helpers (SetRect, Grid, Axes, PointFromIndex,
IndexFromPoint, SetUniformTips, gridScales, corners, shape) and the same
dispatch on 'mode'.  'gpu' returns device arrays (torch tensors on the current
HIP device, the counterpart of the reference's cupy arrays); 'gpu_transfer'
returns numpy arrays.  The CPU modes of the reference call the external
HamiltonFastMarching library, which this package does not provide.

Design: both containers are thin mappings over a plain dict (`store`) that
share one alias rule (a singular key such as 'seed' reads and writes a
one-element list under its plural key).  The grid of a dictIn is described
by three stored keys -- 'dims', 'origin' (lower corner of the box) and
'gridScale' or 'gridScales' -- and every grid helper derives from them.  The
float arithmetic of SetRect and PointFromIndex is carried out in the
container's float type (float32 in the GPU modes) so that dims, origin and
continuous seed indices come out exactly as the reference computes them
(DictIn.py:404-453): the seed at the centre of an even grid lands exactly on
a rounding tie.
"""
from collections.abc import MutableMapping

import numpy as np

# Keys whose values are stored as arrays of the container's float type.
_array_float_fields = {
    'origin', 'dims', 'gridScale', 'gridScales', 'values',
    'seeds', 'seeds_Unoriented', 'tips', 'tips_Unoriented',
    'seedValues', 'seedValues_Unoriented', 'seedValueVariation', 'seedFlags',
    'cost', 'speed', 'costVariation',
    'inspectSensitivity', 'inspectSensitivityWeights', 'inspectSensitivityLengths',
    'exportVoronoiFlags', 'factoringIndexShift',
    'chart_mapping',
    'control', 'controls', 'state_transition_costs',
}

# Singular input keys and the plural key that holds them as a list of one.
_singleIn = {
    'seed': 'seeds', 'seedValue': 'seedValues',
    'seed_Unoriented': 'seeds_Unoriented', 'seedValue_Unoriented': 'seedValues_Unoriented',
    'tip': 'tips', 'tip_Unoriented': 'tips_Unoriented',
    'seedFlag': 'seedFlags', 'seedFlag_Unoriented': 'seedFlags_Unoriented',
}

_readonlyIn = {'float_t'}

# Where a mode keeps its arrays: two modes are interchangeable on a container
# only if they agree here.
_array_module = {
    'cpu': 'host', 'cpu_raw': 'host', 'gpu_transfer': 'host',
    'gpu': 'device', 'cpu_transfer': 'device',
}

# Modes whose default float type is float32 (the GPU arithmetic); the others
# default to float64.
_float32_modes = {'gpu', 'cpu_transfer', 'gpu_transfer'}

_singleOut = {
    'geodesic': 'geodesics', 'geodesic_Unoriented': 'geodesics_Unoriented',
    'geodesic_euclideanLength': 'geodesics_euclideanLength',
}

SEModels = {'ReedsShepp2', 'ReedsSheppForward2', 'Elastica2', 'Dubins2',
            'ReedsSheppExt2', 'ReedsSheppForwardExt2', 'ElasticaExt2', 'ElasticaExt2_5', 'DubinsExt2',
            'ConvexReedsSheppForward2', 'ConvexElastica2', 'ConvexDubins2',
            'ReedsShepp3', 'ReedsSheppForward3'}

dimModels = {'ElasticaExt2_5': 3, 'Riemann3_Periodic': 3, 'ReedsSheppGPU3': 5}

# The keys that describe the grid: SetRect sets three of them at once.
_DOMAIN_KEYS = ('gridScale', 'gridScales', 'dims', 'origin')


def is_tensor(x):
    return type(x).__module__.startswith('torch')


def _host(x):
    """numpy view of an array that may live on the device."""
    return x.detach().cpu().numpy() if is_tensor(x) else np.asarray(x)


class _AliasedStore(MutableMapping):
    """Mapping over `self.store` in which the keys of `_aliases` stand for a
    one-element list kept under their plural key."""

    _aliases = {}

    def _plural(self, key):
        return self._aliases.get(key, key)

    def _put(self, key, value):
        plural = self._plural(key)
        self.store[plural] = [value] if plural != key else value

    def __getitem__(self, key):
        plural = self._plural(key)
        value = self.store[plural]
        if plural == key:
            return value
        if len(value) != 1:
            raise ValueError(f"Found {len(value)} values for key {key}")
        return value[0]

    def __delitem__(self, key):
        del self.store[self._plural(key)]

    def __iter__(self):
        return iter(self.store)

    def __len__(self):
        return len(self.store)

    def keys(self):
        return self.store.keys()

    def __repr__(self):
        return f"{type(self).__name__}({self.store})"


class dictOut(_AliasedStore):
    """Dictionary-like output of the eikonal solver."""

    _aliases = _singleOut

    def __init__(self, store=None):
        self.store = {} if store is None else store

    def __setitem__(self, key, value):
        self._put(key, value)

    def copy(self):
        return dictOut(dict(self.store))

    __copy__ = copy


def CenteredLinspace(a, b, n):
    """The n cell centres of [a, b]: a + (i + 1/2) (b - a)/n.

    Evaluated as (i h + a) + h/2, the order of numpy.linspace(a, b, n,
    endpoint=False) shifted by half its step, so that the centres agree with
    the reference's (DictIn.py:97-107) to the last bit.  n may be given as an
    integral float."""
    count = int(n)
    if count != n or count < 1:
        raise ValueError(f"CenteredLinspace needs a positive integer count, got {n}")
    a = np.asarray(a)
    step = (np.asarray(b) - a) / count
    return np.arange(count, dtype=step.dtype) * step + a + step / 2


class dictIn(_AliasedStore):
    """Dictionary-like input of the eikonal solver (agd/Eikonal/DictIn.py:110).

    Keys in _array_float_fields are cast to arrays of float_t on assignment
    (float32 in the GPU modes); torch tensors stay tensors.
    """

    default_mode = 'gpu'
    _aliases = _singleIn

    def __init__(self, store=None):
        init = dict(store) if store else {}
        mode = init.get('mode', dictIn.default_mode)
        if mode not in _array_module:
            raise ValueError(f"Unknown mode {mode!r}; expected one of {sorted(_array_module)}")
        self._mode = mode
        self._float_t = init.get('float_t', np.float32 if mode in _float32_modes else np.float64)
        self.store = {'arrayOrdering': 'RowMajor'}
        for key in ('mode', 'float_t'):
            if key in init:
                self.store[key] = init[key]
        self.update(init)

    # ------- storage and casting ------
    @property
    def mode(self):
        return self._mode

    @property
    def float_t(self):
        return self._float_t

    def array_float_caster(self, x):
        """x as an array of float_t: torch tensors keep their device, denseAD
        values cast their value and coefficients alike."""
        from ..AutomaticDifferentiation import Dense
        if isinstance(x, Dense.denseAD):
            return Dense.denseAD(self.array_float_caster(x.value), self.array_float_caster(x.coef))
        if is_tensor(x):
            import torch
            return x.to(dtype=torch.float32 if self.float_t == np.float32 else torch.float64)
        return np.asarray(x, dtype=self.float_t)

    def copy(self):
        return dictIn(dict(self.store))

    __copy__ = copy

    def __setitem__(self, key, value):
        if key == 'mode':
            if _array_module[value] != _array_module[self._mode]:
                raise ValueError('Switching between modes with distinct array storage')
            self._mode = value
        if key in _readonlyIn and self.store[key] != value:
            raise ValueError(f"Key {key} is readonly (set at init)")
        plural = self._plural(key)
        if plural != key:
            value = [value]
        if plural in _array_float_fields:
            already = isinstance(value, np.ndarray) and value.dtype == self.float_t
            if not already:
                value = self.array_float_caster(value)
        self.store[plural] = value

    # ------- running the solver ------
    def RunHelp(self, mode=None):
        from . import HFM_GPU
        help(HFM_GPU.RunGPU)

    def Run(self, join=None, **kwargs):
        """Run the eikonal solver and return a dictOut (DictIn.py:248-285)."""
        if join is not None:
            merged = self.copy()
            merged.update(join)
            return merged.Run(**kwargs)
        if self['arrayOrdering'] != 'RowMajor':
            raise ValueError("Unsupported array ordering")
        if self.mode not in ('gpu', 'gpu_transfer'):
            raise NotImplementedError(
                f"mode={self.mode!r} calls the HamiltonFastMarching CPU library, which is not "
                "part of this package; use mode='gpu' or 'gpu_transfer'")
        from . import HFM_GPU
        extra = {'host_outputs': True} if self.mode == 'gpu_transfer' else {}
        return _wrap_outputs(HFM_GPU.RunGPU(self, **extra, **kwargs))

    # ------- grid ------
    @property
    def shape(self):
        return tuple(int(n) for n in _host(self['dims']))

    @property
    def size(self):
        return int(np.prod(self.shape))

    @property
    def SE(self):
        return self['model'] in SEModels

    @property
    def vdim(self):
        model = self['model']
        if model in dimModels:
            return dimModels[model]
        d = int(model[-1])
        return 2 * d - 1 if self.SE else d

    @property
    def nTheta(self):
        """Cells of the angle axis of a curvature-penalised model, covering
        [0, 2 pi) (DictIn.py:323-339); set after SetRect of the physical box."""
        if not self.SE:
            raise ValueError("nTheta belongs to the curvature penalized models")
        shape = self.shape
        if len(shape) != self.vdim:
            raise ValueError("Angular resolution not set")
        return shape[-1]

    @nTheta.setter
    def nTheta(self, value):
        if not self.SE:
            raise ValueError("nTheta belongs to the curvature penalized models")
        if self.vdim != 3:
            raise NotImplementedError("only the 2D curvature penalized models are provided")
        phys = self.shape[:2]
        self['dims'] = self.array_float_caster(list(phys) + [value])

    @property
    def gridScales(self):
        """Cell side along each axis; for a curvature-penalised model
        (gridScale, gridScale, 2 pi / nTheta)."""
        if self.SE:
            h = self['gridScale']
            return self.array_float_caster((h, h, 2. * np.pi / self.nTheta))
        if 'gridScales' in self:
            return self['gridScales']
        return self.array_float_caster((self['gridScale'],) * self.vdim)

    @property
    def corners(self):
        """Lower and upper corners of the box; the angle axis of a
        curvature-penalised model starts half a cell below 0, so that its cell
        centres are the angles 2 pi k / nTheta."""
        h = _host(self.gridScales)
        n = _host(self['dims'])
        nphys = (len(n) + 1) // 2 if self.SE else len(n)
        lo = _host(self['origin']) if 'origin' in self else np.zeros_like(n[:nphys])
        if self.SE and len(lo) < len(n):
            lo = np.concatenate([lo, -h[len(lo):] / 2])
        return lo, lo + h * n

    def Axes(self, dims=None):
        """Cell-centre coordinates along each axis (with `dims` cells per axis
        instead of the grid's own when given)."""
        lo, hi = self.corners
        counts = _host(self['dims']) if dims is None else dims
        return [self.array_float_caster(CenteredLinspace(a, b, n)) for a, b, n in zip(lo, hi, counts)]

    def Grid(self, dims=None):
        """All cell centres, coordinate first: shape (vdim, *dims)."""
        ordering = self['arrayOrdering']
        if ordering not in ('RowMajor', 'YXZ_RowMajor'):
            raise ValueError('Unsupported arrayOrdering : ' + ordering)
        mesh = np.meshgrid(*self.Axes(dims), indexing='ij' if ordering == 'RowMajor' else 'xy',
                           copy=False)
        return np.stack(mesh)

    def SetUniformTips(self, dims):
        """Geodesic tips at the cell centres of a coarser grid `dims`."""
        self['tips'] = self.Grid(dims).reshape(self.vdim, -1).T

    def SetRect(self, sides, sampleBoundary=False, gridScale=None, gridScales=None,
                dimx=None, dims=None):
        """Box domain `sides` = [[lo_0, hi_0], ...] (DictIn.py:404-440).

        The cell side comes from whichever of gridScale, gridScales (per
        axis), dimx (cells along the first axis) or dims (cells per axis) is
        given, in that order of precedence.  Points are cell centres, or lie
        on the box faces with sampleBoundary.  The cell counts are the
        rounded box widths over the cell side, and the box is re-centred on
        the rounded grid."""
        present = sum(key in self for key in _DOMAIN_KEYS)
        if present not in (0, 3):
            raise ValueError("Domain already partially set")
        box = self.array_float_caster(sides)
        lo, hi = box[:, 0], box[:, 1]
        faces = float(sampleBoundary)      # 1: points on the faces, 0: cell centres
        h = self._cell_sides(hi - lo, faces, gridScale, gridScales, dimx, dims)
        # Box width in cells, its rounding, and the half-leftover shift of
        # the lower corner, in float_t and in the reference's order of
        # operations.
        cells = (hi - lo) / h + faces
        counts = np.round(cells)
        if not np.min(counts) > 0:
            raise ValueError(f"SetRect gives an empty grid: {counts}")
        self['dims'] = counts
        self['origin'] = lo + (cells - counts - faces) * h / 2

    def _cell_sides(self, width, faces, gridScale, gridScales, dimx, dims):
        """Stores the grid-scale key SetRect derives from its arguments and
        returns the per-axis cell side (float_t)."""
        ndim = len(width)
        cast = self.array_float_caster
        if gridScale is not None:
            self['gridScale'] = gridScale
            return cast([gridScale] * ndim)
        if gridScales is not None:
            self['gridScales'] = gridScales
            return cast(gridScales)
        if dimx is not None:
            side = width[0] / (dimx - faces)
            self['gridScale'] = side
            return cast([side] * ndim)
        if dims is not None:
            sides = width / (cast(dims) - faces)
            self['gridScales'] = sides
            return cast(sides)
        raise ValueError('Missing argument gridScale, gridScales, dimx, or dims')

    def PointFromIndex(self, index, to=False):
        """Cell index -> point x = lo + h/2 + h index; with to=True, point ->
        continuous index (x - (lo + h/2)) / h (DictIn.py:442-453).  Float
        arithmetic in float_t."""
        lo, _ = self.corners
        h = _host(self.gridScales)
        first = lo + self.float_t(0.5) * h      # centre of cell 0
        x = np.asarray(_host(index), dtype=self.float_t)
        if x.shape[-1] != self.vdim:
            raise ValueError(f"points of dimension {self.vdim} expected, got shape {x.shape}")
        return (x - first) / h if to else first + h * x

    def OrientedPoints(self, pointU):
        """Each physical point paired with every angle of the grid: shape
        (nTheta, ..., 3) (DictIn.py:464-482)."""
        if not self.SE:
            raise ValueError("OrientedPoints only makes sense for models SE space.")
        if self.vdim != 3:
            raise ValueError("Sorry, oriented point not implemented for SE(3) models.")
        p = np.asarray(_host(pointU), dtype=self.float_t)
        if p.shape[-1] != 2:
            raise ValueError(f"Unoriented points expected to have 2 dimensions, found {p.shape[-1]}.")
        theta = _host(self.Axes()[2])
        out = np.empty((len(theta),) + p.shape[:-1] + (3,), dtype=self.float_t)
        out[..., :2] = p
        out[..., 2] = np.asarray(theta, dtype=self.float_t).reshape((-1,) + (1,) * (p.ndim - 1))
        return out

    def IndexFromPoint(self, point):
        """Nearest cell index of each point (ties to even) and the offset of
        the point from that cell's centre, in cells."""
        cont = self.PointFromIndex(point, to=True)
        nearest = np.round(cont)
        return nearest.astype(int), cont - nearest


def _wrap_outputs(out):
    """The solver's dict outputs as dictOut, within a tuple/list return."""
    if isinstance(out, dict):
        return dictOut(out)
    if isinstance(out, (tuple, list)):
        return type(out)(dictOut(o) if isinstance(o, dict) else o for o in out)
    return out
