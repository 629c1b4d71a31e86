"""dictIn / dictOut: the input and output containers of the eikonal solver.

Drop-in counterpart of agd/Eikonal/DictIn.py for the GPU path: same keys,
same grid helpers (SetRect, Grid, Axes, PointFromIndex, IndexFromPoint,
SetUniformTips, gridScales, corners, shape), same dispatch on 'mode'.
'gpu' returns device arrays (torch tensors on the current HIP device, the
counterpart of the reference's cupy arrays); 'gpu_transfer' returns numpy.
The CPU modes of the reference call the external HamiltonFastMarching
library, which this package does not provide.
"""
from collections.abc import MutableMapping

import numpy as np

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

_singleIn = {
    'seed': 'seeds', 'seedValue': 'seedValues',
    'seed_Unoriented': 'seeds_Unoriented', 'seedValue_Unoriented': 'seedValues_Unoriented',
    'tip': 'tips', 'tip_Unoriented': 'tips_Unoriented',
    'seedFlag': 'seedFlags', 'seedFlag_Unoriented': 'seedFlags_Unoriented',
}

_readonlyIn = {'float_t'}

_array_module = {
    'cpu': 'host', 'cpu_raw': 'host', 'gpu_transfer': 'host',
    'gpu': 'device', 'cpu_transfer': 'device',
}

_singleOut = {
    'geodesic': 'geodesics', 'geodesic_Unoriented': 'geodesics_Unoriented',
    'geodesic_euclideanLength': 'geodesics_euclideanLength',
}

SEModels = {'ReedsShepp2', 'ReedsSheppForward2', 'Elastica2', 'Dubins2',
            'ReedsSheppExt2', 'ReedsSheppForwardExt2', 'ElasticaExt2', 'ElasticaExt2_5', 'DubinsExt2',
            'ConvexReedsSheppForward2', 'ConvexElastica2', 'ConvexDubins2',
            'ReedsShepp3', 'ReedsSheppForward3'}

dimModels = {'ElasticaExt2_5': 3, 'Riemann3_Periodic': 3, 'ReedsSheppGPU3': 5}


def is_tensor(x):
    return type(x).__module__.startswith('torch')


class dictOut(MutableMapping):
    """Dictionary-like output of the eikonal solver."""

    def __init__(self, store=None):
        self.store = store if store is not None else {}

    def __copy__(self):
        return dictOut(self.store.copy())

    def copy(self):
        return self.__copy__()

    def __repr__(self):
        return f"dictOut({self.store})"

    def __setitem__(self, key, value):
        if key in _singleOut:
            key = _singleOut[key]
            value = [value]
        self.store[key] = value

    def __getitem__(self, key):
        if key in _singleOut:
            values = self.store[_singleOut[key]]
            if len(values) != 1:
                raise ValueError(f"Found {len(values)} values for key {key}")
            return values[0]
        return self.store[key]

    def __delitem__(self, key):
        del self.store[_singleOut.get(key, key)]

    def __iter__(self):
        return iter(self.store)

    def __len__(self):
        return len(self.store)

    def keys(self):
        return self.store.keys()


def CenteredLinspace(a, b, n):
    """linspace shifted by half a step (cell centres of [a,b] split in n)."""
    n_ = int(n)
    assert n == n_
    r, dr = np.linspace(a, b, n_, endpoint=False, retstep=True)
    if np.any(np.isnan(dr)):
        assert n == 1
        dr = b - a
    return r + dr / 2


class dictIn(MutableMapping):
    """Dictionary-like input of the eikonal solver (agd/Eikonal/DictIn.py:110).

    Keys in _array_float_fields are cast to arrays of float_t on assignment
    (float32 in the GPU modes); torch tensors are kept as tensors.
    """

    default_mode = 'gpu'

    def __init__(self, store=None):
        if store is None:
            store = dict()
        self.store = {'arrayOrdering': 'RowMajor'}
        if 'mode' in store:
            mode = store['mode']
            self.store['mode'] = mode
        else:
            mode = dictIn.default_mode
        assert mode in ('cpu', 'cpu_raw', 'cpu_transfer', 'gpu', 'gpu_transfer')
        self._mode = mode
        float_t = np.float32 if mode in ('gpu', 'cpu_transfer', 'gpu_transfer') else np.float64
        if 'float_t' in store:
            float_t = store['float_t']
            self.store['float_t'] = float_t
        self._float_t = float_t
        if store:
            self.update(store)

    def array_float_caster(self, x):
        from ..AutomaticDifferentiation import Dense
        if isinstance(x, Dense.denseAD):  # cast the value and the coefficients alike
            return Dense.denseAD(self.array_float_caster(x.value), self.array_float_caster(x.coef))
        if is_tensor(x):
            import torch
            return x.to(dtype=torch.float32 if self.float_t == np.float32 else torch.float64)
        return np.asarray(x, dtype=self.float_t)

    def __copy__(self):
        return dictIn(self.store.copy())

    def copy(self):
        return self.__copy__()

    @property
    def mode(self):
        return self._mode

    @property
    def float_t(self):
        return self._float_t

    def __repr__(self):
        return f"dictIn({self.store})"

    def __setitem__(self, key, value):
        if key == 'mode':
            if _array_module[value] != _array_module[self.mode]:
                raise ValueError('Switching between modes with distinct array storage')
            self._mode = value
        if key in _readonlyIn and self.store[key] != value:
            raise ValueError(f"Key {key} is readonly (set at init)")
        if key in _singleIn:
            key = _singleIn[key]
            value = [value]
        if key in _array_float_fields and not (isinstance(value, np.ndarray) and
                                               value.dtype == self.float_t):
            value = self.array_float_caster(value)
        self.store[key] = value

    def __getitem__(self, key):
        if key in _singleIn:
            values = self.store[_singleIn[key]]
            if len(values) != 1:
                raise ValueError(f"Found {len(values)} values for key {key}")
            return values[0]
        return self.store[key]

    def __delitem__(self, key):
        del self.store[_singleIn.get(key, key)]

    def __iter__(self):
        return iter(self.store)

    def __len__(self):
        return len(self.store)

    def keys(self):
        return self.store.keys()

    def RunHelp(self, mode=None):
        from . import HFM_GPU
        help(HFM_GPU.RunGPU)

    def Run(self, join=None, **kwargs):
        """Run the eikonal solver and return a dictOut (DictIn.py:248-285)."""
        if join is not None:
            other = self.copy()
            other.update(join)
            return other.Run(**kwargs)
        if self['arrayOrdering'] != 'RowMajor':
            raise ValueError("Unsupported array ordering")

        def to_dictOut(out):
            if isinstance(out, (tuple, list)) and len(out) and isinstance(out[0], dict):
                return (dictOut(out[0]),) + tuple(out[1:])
            if isinstance(out, (tuple, list)):
                return type(out)(dictOut(o) if isinstance(o, dict) else o for o in out)
            return dictOut(out)

        if self.mode in ('cpu', 'cpu_raw', 'cpu_transfer'):
            raise NotImplementedError(
                f"mode={self.mode!r} calls the HamiltonFastMarching CPU library, which is not "
                "part of this package; use mode='gpu' or 'gpu_transfer'")
        from . import HFM_GPU
        if self.mode == 'gpu':
            return to_dictOut(HFM_GPU.RunGPU(self, **kwargs))
        out = HFM_GPU.RunGPU(self, host_outputs=True, **kwargs)
        return to_dictOut(out)

    # ------- Grid related functions ------
    @property
    def shape(self):
        dims = self['dims']
        if is_tensor(dims):
            dims = dims.cpu().numpy()
        return tuple(np.asarray(dims).astype(int))

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
        dim = int(model[-1])
        return (2 * dim - 1) if self.SE else dim

    @property
    def gridScales(self):
        if self.SE:
            raise NotImplementedError("Curvature penalized models are not provided")
        if 'gridScales' in self:
            return self['gridScales']
        return self.array_float_caster((self['gridScale'],) * self.vdim)

    @property
    def corners(self):
        gridScales = _host(self.gridScales)
        dims = _host(self['dims'])
        origin = _host(self.get('origin', np.zeros_like(dims)))
        return (origin, origin + gridScales * dims)

    def Axes(self, dims=None):
        bottom, top = self.corners
        if dims is None:
            dims = _host(self['dims'])
        return [self.array_float_caster(CenteredLinspace(b, t, d)) for b, t, d in zip(bottom, top, dims)]

    def Grid(self, dims=None):
        axes = self.Axes(dims)
        ordering = self['arrayOrdering']
        if ordering == 'RowMajor':
            return np.array(np.meshgrid(*axes, indexing='ij', copy=False))
        if ordering == 'YXZ_RowMajor':
            return np.array(np.meshgrid(*axes, copy=False))
        raise ValueError('Unsupported arrayOrdering : ' + ordering)

    def SetUniformTips(self, dims):
        self['tips'] = self.Grid(dims).reshape(self.vdim, -1).T

    def SetRect(self, sides, sampleBoundary=False, gridScale=None, gridScales=None,
                dimx=None, dims=None):
        """Box domain (DictIn.py:404-440): cell-centred sampling unless
        sampleBoundary, grid scale from gridScale(s), dimx or dims."""
        domain_count = sum(e in self for e in ('gridScale', 'gridScales', 'dims', 'origin'))
        if domain_count not in (0, 3):
            raise ValueError("Domain already partially set")
        caster = self.array_float_caster
        corner0, corner1 = caster(sides).T
        dim = len(corner0)
        sb = float(sampleBoundary)
        width = corner1 - corner0
        if gridScale is not None:
            gridScales = [gridScale] * dim
            self['gridScale'] = gridScale
        elif gridScales is not None:
            self['gridScales'] = gridScales
        elif dimx is not None:
            gridScale = width[0] / (dimx - sb)
            gridScales = [gridScale] * dim
            self['gridScale'] = gridScale
        elif dims is not None:
            gridScales = width / (caster(dims) - sb)
            self['gridScales'] = gridScales
        else:
            raise ValueError('Missing argument gridScale, gridScales, dimx, or dims')
        h = caster(gridScales)
        ratios = (corner1 - corner0) / h + sb
        dims = np.round(ratios)
        assert np.min(dims) > 0
        origin = corner0 + (ratios - dims - sb) * h / 2
        self['dims'] = dims
        self['origin'] = origin

    def PointFromIndex(self, index, to=False):
        """Index -> point; with to=True, point -> continuous index
        (DictIn.py:442-453).  Arithmetic in float_t, as the reference."""
        bottom, _ = self.corners
        scale = _host(self.gridScales)
        start = bottom + self.float_t(0.5) * scale
        index = np.asarray(_host(index), dtype=self.float_t)
        assert index.shape[-1] == self.vdim
        if not to:
            return start + scale * index
        return (index - start) / scale

    def IndexFromPoint(self, point):
        continuousIndex = self.PointFromIndex(point, to=True)
        index = np.round(continuousIndex)
        return index.astype(int), (continuousIndex - index)


def _host(x):
    if is_tensor(x):
        return x.detach().cpu().numpy()
    return np.asarray(x)
