"""GPU eikonal solver driver: the counterpart of agd/Eikonal/HFM_CUDA.

RunGPU(hfmIn) reads the dictIn keys (with defaults and help, as
_Interface.py:80-122), prepares the tiled device arrays, calls the HIP
library (libagd_hip.so) and assembles hfmOut.  Pipeline order follows
_Interface.py:128-146:
    SetKernelTraits -> SetGeometry -> SetArgs -> Solve('eikonal')
    -> PostProcess (flow) -> SolveAD (reverse) -> FinalCheck.

Device buffers are torch tensors on the current HIP device (allocation and
stream plumbing only); all computation happens in the HIP kernels.
"""
import ctypes
import numbers
import time
from collections import OrderedDict

import numpy as np

from .. import AutomaticDifferentiation as ad
from .. import FiniteDifferences as fd
from .. import Metrics
from .. import _lib
from . import _factor
from ._device import (torch, device_tensor, host_array, ptr, stream_handle, empty,
                      full, zeros)

_MODELS = {'Isotropic': _lib.ISOTROPIC, 'Diagonal': _lib.DIAGONAL,
           'Riemann': _lib.RIEMANN, 'Rander': _lib.RANDER, 'TTI': _lib.TTI,
           'ReedsShepp': _lib.REEDSSHEPP2, 'ReedsSheppForward': _lib.REEDSSHEPPFWD2,
           'Elastica': _lib.ELASTICA2, 'Dubins': _lib.DUBINS2}
# Curvature-penalised models: 2D position and an angle, on (x, y, theta) grids.
_CURVATURE = ('ReedsShepp', 'ReedsSheppForward', 'Elastica', 'Dubins')


def scheme_slots(model_, ndim):
    """Offsets per (sub-)scheme of a model, each stored as a pair of
    neighbours (the forward offsets of the curvature models included)."""
    if model_ in ('Isotropic', 'Diagonal'):
        return ndim
    return {'ReedsShepp': 6, 'ReedsSheppForward': 7, 'Elastica': 30, 'Dubins': 6}.get(
        model_, ndim * (ndim + 1) // 2)

_SOLVER_NAMES = {'AGSI': 'adaptive_gauss_siedel_iteration', 'FIM': 'fast_iterative_method'}
_SOLVER_IDS = {'adaptive_gauss_siedel_iteration': _lib.AGSI,
               'fast_iterative_method': _lib.FIM, 'global_iteration': _lib.GLOBAL}


def default_traits(model):
    """Tile shape, in-tile sweeps and FIM front width per model
    (kernel_traits.py:6-137)."""
    table = {
        'Isotropic2': ((16, 16), 16, 4), 'Diagonal2': ((16, 16), 16, 4),
        'Isotropic3': ((4, 4, 4), 8, 4), 'Diagonal3': ((4, 4, 4), 8, 4),
        'Riemann2': ((8, 8), 8, 4), 'Riemann3': ((4, 4, 4), 4, 6),
        'Rander2': ((8, 8), 12, 8), 'Rander3': ((4, 4, 4), 8, 8),
        'TTI2': ((8, 8), 10, 6), 'TTI3': ((4, 4, 4), 3, 6),
        # curvature models on (x, y, theta) grids (kernel_traits.py:23-37; the
        # 4x4x2 tiles there are 4x4x4 here, the smallest 3D tile).  The two
        # Reeds-Shepp models take 2 sweeps and FIM (front width 5) instead of
        # the reference's 6 sweeps: measured on 192 x 192 x 64, 4.91 -> 3.64 ms
        # (ReedsShepp2, AGSI before) and 4.17 -> 3.64 ms (ReedsSheppForward2),
        # the same fixed point; Dubins2 takes FIM (front width 4) instead of
        # AGSI: 6.44 -> 5.22 ms
        'ReedsShepp2': ((4, 4, 4), 2, 5), 'ReedsSheppForward2': ((4, 4, 4), 2, 5),
        'Elastica2': ((4, 4, 4), 1, 4), 'Dubins2': ((4, 4, 4), 1, 4),
    }
    if model not in table:
        raise NotImplementedError(
            f"Model {model} is not provided by the MI355X eikonal solver "
            f"(supported: {sorted(table)})")
    shape_i, niter_i, fim = table[model]
    traits = {'Scalar': np.float32, 'Int': np.int32, 'OffsetT': np.int32,
              'shape_i': shape_i, 'niter_i': niter_i}
    return traits, fim


def RunGPU(hfmIn, *args, cache=None, host_outputs=False, **kwargs):
    """Run the GPU eikonal solver (HFM_CUDA/__init__.py:10-25).

    Set hfmIn['verbosity'] (0, 1, 2) to control the printed information;
    hfmOut['keys']['help'] documents every key read.
    """
    if cache is not None:
        print("Warning : gpu eikonal solver does not support caching")
    return Interface(hfmIn, host_outputs=host_outputs).Run(*args, **kwargs)


class Interface:
    """Carries out RunGPU (_Interface.py:23-174).  Not meant for direct use."""

    def __init__(self, hfmIn, host_outputs=False, batched=False):
        self.hfmIn = hfmIn
        self.batched = batched  # member of a batched solve (tiles below 64 cells allowed)
        if hfmIn['arrayOrdering'] != 'RowMajor':
            raise ValueError("Only RowMajor indexing supported")
        self.host_outputs = host_outputs
        self.hfmOut = {'keys': {
            'used': ['origin', 'arrayOrdering', 'dims', 'mode', 'projective'],
            'default': OrderedDict(), 'visited': [], 'help': OrderedDict(),
            'kernelStats': OrderedDict()}}
        self.verbosity = 1
        self.verbosity = self.GetValue('verbosity', default=1,
                                       help="Choose the amount of detail displayed on the run")
        self.model = self.GetValue('model', help='Minimal path model to be solved.')
        # _Interface.py:51-55: the extended curvature models are the standard
        # ones; the Convex prefix is a variant of the scheme
        if self.model == 'ElasticaExt2_5':
            self.model = 'ElasticaExt2'
        if self.model.endswith('Ext2'):
            self.model = self.model[:-4] + '2'
        self.convex = self.model.startswith('Convex')
        if self.convex:
            self.model = self.model[6:]
        self.ndim = len(hfmIn['dims'])
        self.stats = {key: {} for key in ('eikonal', 'flow', 'reverseAD')}

    @property
    def model_(self):
        return self.model[:-1]

    def HasValue(self, key):
        self.hfmOut['keys']['visited'].append(key)
        return key in self.hfmIn

    def GetValue(self, key, default="_None", verbosity=2, help="Sorry : no help for this key"):
        """Value of a key, recording its default and help (_Interface.py:80-122)."""
        assert key not in self.hfmOut['keys']['help']
        self.hfmOut['keys']['help'][key] = help
        self.hfmOut['keys']['default'][key] = default
        if key in self.hfmIn:
            self.hfmOut['keys']['used'].append(key)
            return self.hfmIn[key]
        if isinstance(default, str) and default == "_None":
            raise ValueError(f"Missing value for key {key}")
        if verbosity <= self.verbosity:
            if isinstance(default, str) and default == "_Dummy":
                print(f"see out['keys']['default'][{key}] for default")
            else:
                print(f"key {key} defaults to {default}")
        return default

    def Warn(self, msg):
        if self.verbosity >= -1:
            print("---- Warning ----\n", msg, "\n-----------------\n")

    def Run(self):
        self.SetKernelTraits()
        self.SetGeometry()
        self.SetArgs()
        self.Solve()
        self.PostProcess()
        self.SolveAD()
        self.GetGeodesics()
        self.FinalCheck()
        if self.extractValues or self.retself:
            retval = [self.hfmOut]
            if self.extractValues:
                retval.insert(0, self.values)
            if self.retself:
                retval.append(self)
            return retval
        return self.hfmOut

    # ------------------------------------------------------------ traits
    def SetKernelTraits(self):
        """_Kernel.py:23-92."""
        if self.verbosity >= 1:
            print("Setting the kernel traits.")
        self.curvature = self.model_ in _CURVATURE
        if self.convex and self.model not in ('ReedsSheppForward2', 'Elastica2', 'Dubins2'):
            raise ValueError(f"Unsupported model Convex{self.model}")
        phys = int(self.model[-1]) if self.model[-1].isdigit() else -1
        if self.model_ not in _MODELS or self.ndim not in (2, 3) or \
                (phys != self.ndim and not (self.curvature and phys == 2 and self.ndim == 3)):
            default_traits(self.model)  # raises with the list of models
        traits, fim_default = default_traits(self.model)
        user = self.GetValue('traits', default=traits,
                             help="Optional trait parameters for the eikonal kernel.")
        traits = {**traits, **(user or {})}
        self.traits = traits
        # Multiprecision (_Kernel.py:36-44): the reference splits values into
        # float32 + integer multiples of a step; here the solve runs in FP64
        # arithmetic (same rate as FP32 on the vector units), with the
        # reference's float32-based default tolerances and rtol = 0.
        self.values_float64 = bool(self.GetValue(
            'values_float64', default=False,
            help="Export values in float64 (implies multiprecision)"))
        self.multiprecision = self.values_float64 or bool(self.GetValue(
            'multiprecision', default=False,
            help="Use multiprecision arithmetic, to improve accuracy (FP64 arithmetic here)"))
        # Source factorization (_Kernel.py:47-49, 301-308): around a single
        # seed, neighbour values are shifted by the finite-difference error of
        # the distance for the metric frozen at the seed.
        self.factoringRadius = float(self.GetValue(
            'factoringRadius', default=0,
            help="Use source factorization, to improve accuracy"))
        if self.factoringRadius < 0:
            raise ValueError("factoringRadius must be >= 0")
        if self.factoringRadius:
            # The point choice concerns the CPU solver's factorization; the
            # GPU kernel always freezes the metric at the seed.
            self.GetValue('factoringPointChoice', default='Seed',
                          help="Where the factorization metric is taken (the seed, on the GPU)")
        self.order = int(self.GetValue('order', default=1,
                                       help="Scheme order (1, or 2: second order where the solution is smooth)"))
        if self.order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        if self.order == 2:
            self.order2_threshold = float(self.GetValue(
                'order2_threshold', default=0.3,
                help="Relative threshold on second order differences, for the second-order scheme"))
        # _Kernel.py:64-69, _SetGeometry.py:36-37: freeze the tiles whose
        # smallest change lies above a moving threshold (AGSI only).
        self.bound_active_blocks = self.GetValue(
            'bound_active_blocks', default=False,
            help="Limit the number of active blocks in the front. "
                 "Admissible values : (False,True, or positive integer)")

        self.float_t = np.dtype(traits['Scalar']).type
        if self.float_t not in (np.float32, np.float64):
            raise ValueError(f"Unsupported Scalar trait {traits['Scalar']}")
        self.tol_float_t = self.float_t      # resolution behind the default tolerances
        self.export_t = np.float64 if self.values_float64 else self.float_t
        if self.multiprecision:
            self.float_t = np.float64
        self.scalar = _lib.F32 if self.float_t == np.float32 else _lib.F64
        self.shape_i = tuple(int(s) for s in traits['shape_i'])
        self.niter_i = int(traits['niter_i'])
        if len(self.shape_i) != self.ndim:
            raise ValueError(f"shape_i {self.shape_i} does not match dimension {self.ndim}")
        supported = _lib.lib.agd_batch_tile_supported if self.batched else _lib.lib.agd_tile_supported
        if not supported(_MODELS[self.model_], self.ndim, _lib.i32(self.shape_i)):
            raise ValueError(f"No kernel compiled for tile shape {self.shape_i}")
        self.size_i = int(np.prod(self.shape_i))

        solver = 'FIM' if ((self.HasValue('fim_front_width') or fim_default)
                           and not self.bound_active_blocks) else 'AGSI'
        solver = self.GetValue('solver', default=solver,
                               help="Choice of fixed point solver (AGSI, FIM, global_iteration)")
        self.solver = _SOLVER_NAMES.get(solver, solver)
        if self.solver not in _SOLVER_IDS:
            raise ValueError(f"Unrecognized solver : {solver}")
        if self.bound_active_blocks and self.solver != 'adaptive_gauss_siedel_iteration':
            # _Kernel.py:78-80 (the freezing runs in the AGSI loop only)
            raise ValueError(f"Incompatible options found for {self.solver} solver (bound_active_blocks)")
        self.fim_front_width = fim_default
        if self.solver == 'fast_iterative_method':
            self.fim_front_width = self.GetValue(
                'fim_front_width', default=fim_default,
                help="Dictates the max front width in the FIM variant.\n(original FIM : 2. Must be >=2.)")
            if self.fim_front_width is None or self.fim_front_width < 2:
                raise ValueError("Please specify an fim_front_width (integer >=2)")

    # ------------------------------------------------------------ geometry
    def SetGeometry(self):
        """_SetGeometry.py:16-282 (metric, cost, tolerances, walls)."""
        if self.verbosity >= 1:
            print("Preparing the domain data (shape,metric,...)")
        self.shape = self.hfmIn.shape
        self.periodic = tuple(bool(p) for p in self.GetValue(
            'periodic', default=(False, False, True) if self.curvature else (False,) * self.ndim,
            help="Apply periodic boundary conditions on some axes"))
        self.shape_o = tuple(int(s) for s in fd.round_up_ratio(self.shape, self.shape_i))
        self.ntiles = int(np.prod(self.shape_o))
        if self.bound_active_blocks is True:   # _SetGeometry.py:36-37
            self.bound_active_blocks = 12 * np.prod(self.shape_o) / np.max(self.shape_o)
        self.bound_active_blocks = int(np.ceil(self.bound_active_blocks or 0))
        self.size_tot = self.ntiles * self.size_i
        if 'gridScale' in self.hfmIn:
            self.GetValue('gridScale', help="Scale of the computational grid")
        else:
            self.GetValue('gridScales', help="Axis independent scales of the computational grid")
        self.h = np.asarray(host_array(self.hfmIn.gridScales), dtype=np.float32)
        self.stream = stream_handle()
        self.modelId = _MODELS[self.model_]

        # Metric -> tiled geometry (computed on the device).
        self.iso_weights = (0., 0., 0.)
        self.geom = None
        if self.model_ == 'Isotropic':
            w = (self.h.astype(np.float32) ** -2).astype(np.float64)
            self.iso_weights = tuple(w.tolist()) + (0.,) * (3 - self.ndim)
        elif self.curvature:
            self._set_curvature_geometry()
        else:
            metric = self.GetValue('metric', default=None, verbosity=3,
                                   help="Metric of the minimal path model")
            dualMetric = self.GetValue('dualMetric', default=None, verbosity=3,
                                       help="Dual metric of the minimal path model")
            self._set_metric_geometry(metric, dualMetric)

        self.exportGeodesicFlow = self.GetValue(
            'exportGeodesicFlow', default=False,
            help="Export the upwind geodesic flow (direction of the geodesics)")
        tips = self.GetValue('tips', default=None, help="Tips from which to compute the minimal geodesics")
        self.tips = None if tips is None else np.asarray(host_array(tips), dtype=self.float_t).reshape(-1, self.ndim)
        self.tips_Unoriented = None
        if self.curvature:
            tipsU = self.GetValue('tips_Unoriented', default=None,
                                  help="Compute a geodesic from the most favorable orientation")
            if tipsU is not None:
                self.tips_Unoriented = np.asarray(host_array(tipsU), dtype=self.float_t).reshape(-1, 2)
        self.hasTips = self.tips is not None or self.tips_Unoriented is not None

        # Cost function (_SetGeometry.py:221-233): a denseAD cost, or a
        # separate costVariation, requests forward AD.
        if self.HasValue('speed'):
            speed = self.GetValue('speed', help="speed = 1/cost (scales the metric, accepts AD)")
            if ad.is_ad(speed):
                # cost = 1/speed: d cost = -d speed / speed^2 (chain rule of the
                # reference's denseAD division)
                sv = speed.value
                sq = sv * sv
                self.cost = ad.Dense.denseAD(1. / sv, -speed.coef / sq.reshape(tuple(sq.shape) + (1,)))
            else:
                self.cost = 1. / speed
        else:
            self.cost = self.GetValue('cost', default=None, help="cost = 1/speed (scales the metric, accepts AD)")
            if self.cost is None:
                self.cost = np.ones((), dtype=np.float32)
        self.cost_coef = None
        if ad.is_ad(self.cost):
            self.cost, self.cost_coef = self.cost.value, self.cost.coef
        else:
            self.cost_coef = self.GetValue('costVariation', default=None,
                                           help="First order variation of the cost function")
        if self.curvature:
            # a physical (nx, ny) or angular (ntheta) cost field is broadcast
            # over the grid (_Interface.py:221-234); the curvature schemes are
            # in units of the physical cell side h (_SetGeometry.py:234)
            if np.ndim(host_array(self.cost)) > 0:
                self.cost = np.ascontiguousarray(self._angular_field(self.cost, 'cost'))
            self.cost = self.cost * self.curvature_cost_scale
            if self.cost_coef is not None:
                self.cost_coef = self.cost_coef * self.curvature_cost_scale

        # Tolerances (_SetGeometry.py:237-253)
        res = np.finfo(self.tol_float_t).resolution
        if self.HasValue('atol') and self.HasValue('rtol'):
            tol = None
        else:
            tol = self.GetValue('tol', default="_Dummy",
                                help="Convergence tolerance for the fixed point solver (determines atol, rtol)")
            if isinstance(tol, str) and tol == "_Dummy":
                c = self.cost
                mean_cost = float(torch.nanmean(c.float()).item()) if hasattr(c, 'detach') \
                    else float(np.nanmean(np.asarray(c, dtype=np.float64)))
                tol = mean_cost * res * 5.
                self.hfmOut['keys']['default']['tol'] = self.tol_float_t(tol)
        self.atol = float(self.GetValue('atol', default=tol,
                                        help="Absolute convergence tolerance for the fixed point solver"))
        self.rtol = float(self.GetValue('rtol', default=0. if self.multiprecision else float(res * 5.),
                                        help="Relative convergence tolerance for the fixed point solver"))

        # Walls
        self.walls = self.GetValue('walls', default=None, help='Obstacles in the domain')
        if self.walls is not None and self.curvature:
            # physical (nx, ny) or angular walls are broadcast (_SetGeometry.py:265)
            self.walls = np.ascontiguousarray(self._angular_field(self.walls, 'walls')).astype(np.uint8)
        self.wallDistBound = int(self.GetValue(
            'wallDistBound', default=10,
            help="Bound on the computed distance to the obstacles.\n(Ideally a sharp upper bound on the stencil width.)"))
        if self.bound_active_blocks:   # _SetGeometry.py:255-258
            self.minChg_delta_min = float(self.GetValue(
                'minChg_delta_min', default=float(np.min(self.h)) / 10.,
                help="Minimal threshold increase with bound_active_blocks method"))

    def _angular_field(self, value, name):
        """A scalar, a physical field (nx, ny), an angular field (ntheta) or a
        full field, broadcast to the (x, y, theta) grid (_Interface.py:221-236)."""
        a = np.asarray(host_array(value), dtype=np.float64)
        shape = tuple(self.shape)
        if a.ndim == 0 or a.shape == shape:
            return np.broadcast_to(a, shape)
        if a.shape == shape[2:]:
            return np.broadcast_to(a.reshape((1, 1) + shape[2:]), shape)
        if a.shape == shape[:2]:
            return np.broadcast_to(a.reshape(shape[:2] + (1,)), shape)
        raise ValueError(f"Field {name} has incorrect dimensions {a.shape} for the grid {shape}")

    def _set_curvature_geometry(self):
        """Curvature models (_SetGeometry.py:83-109, _Kernel.py:318-332): per
        cell [1/xi, kappa, cos theta, sin theta] in grid units, with 1/xi and
        kappa divided by h_theta / h (the angular over the physical cell
        side); theta defaults to the angle of the grid's third axis; the cost
        is multiplied by h (_SetGeometry.py:234)."""
        h_base, h_per = float(self.h[0]), float(self.h[2])
        ratio = h_per / h_base
        xi = self.GetValue('xi', help="Cost of rotation for the curvature penalized models")
        kappa = self.GetValue('kappa', default=0., help="Rotation bias for the curvature penalized models")
        theta = self.GetValue('theta', default=0., verbosity=3,
                              help="Deviation from horizontality, for the curvature penalized models")
        eps = float(self.GetValue('eps', default=0.1,
                                  help='Relaxation parameter for the curvature penalized models'))
        if not eps > 0:
            raise ValueError("eps must be > 0")
        self.curvature_relax = eps * eps
        ixi = 1. / (self._angular_field(xi, 'xi') * ratio)
        kap = self._angular_field(kappa, 'kappa') / ratio
        if np.ndim(host_array(theta)) == 0:
            # a scalar theta selects the angles of the grid's third axis
            th = self._angular_field(np.asarray(host_array(self.hfmIn.Axes()[2]), dtype=np.float64), 'theta')
        else:
            th = self._angular_field(theta, 'theta')
        geom = np.stack([ixi, kap, np.cos(th), np.sin(th)])
        g_d = device_tensor(np.ascontiguousarray(geom), self.float_t)
        self.geom = empty((4, self.size_tot), self.float_t)
        esize = np.dtype(self.float_t).itemsize
        _lib.check(_lib.lib.agd_block_expand(self.ndim, _lib.i64(self.shape), _lib.i32(self.shape_i), 4,
                                             esize, ptr(g_d), 1, ptr(self.geom),
                                             0., self.stream))
        self._keep = (g_d,)
        self.curvature_cost_scale = h_base

    def _set_metric_geometry(self, metric, dualMetric):
        model = self.model_
        if model == 'TTI':
            return self._set_tti_geometry(metric, dualMetric)
        cls = {'Diagonal': Metrics.Diagonal, 'Riemann': Metrics.Riemann, 'Rander': Metrics.Rander}[model]
        if metric is None and dualMetric is None:
            raise ValueError(f"Missing value for key metric (model {self.model})")
        if metric is not None and not isinstance(metric, Metrics.Base):
            metric = cls.from_HFM(host_array(metric))
        if metric is None:
            if not isinstance(dualMetric, Metrics.Base):
                dualMetric = cls.from_HFM(host_array(dualMetric))
            metric = dualMetric.dual()
        if model == 'Diagonal':
            m = metric.costs if isinstance(metric, Metrics.Diagonal) else None
            if m is None:
                raise ValueError("Diagonal model expects a Metrics.Diagonal metric")
            arr, w = m, None
            core = (self.ndim,)
        elif model == 'Riemann':
            metric = metric if isinstance(metric, Metrics.Riemann) else Metrics.Riemann.from_cast(metric)
            arr, w = metric.m, None
            core = (self.ndim, self.ndim)
        else:
            metric = metric if isinstance(metric, Metrics.Rander) else Metrics.Rander.from_cast(metric)
            arr, w = metric.m, metric.w
            core = (self.ndim, self.ndim)
        self.metric = metric
        arr_d, stride = self._field_tensor(arr, core, 'metric')
        w_d = None
        if w is not None:
            w_d, wstride = self._field_tensor(w, (self.ndim,), 'metric drift')
            if wstride != stride:  # broadcast the constant one
                if stride == 1:
                    w_d = w_d.reshape(self.ndim, 1).expand(self.ndim, int(np.prod(self.shape))).contiguous()
                else:
                    arr_d = arr_d.reshape(-1, 1).expand(arr_d.numel(), int(np.prod(self.shape))).contiguous()
                    stride = 1
        ng = {'Diagonal': self.ndim, 'Riemann': self.ndim * (self.ndim + 1) // 2,
              'Rander': self.ndim * (self.ndim + 1) // 2 + self.ndim}[model]
        self.geom = empty((ng, self.size_tot), self.float_t)
        h = (ctypes_double3(self.h))
        _lib.check(_lib.lib.agd_prepare_geometry(
            self.modelId, self.ndim, _lib.i64(self.shape), _lib.i32(self.shape_i), self.scalar,
            h, ptr(arr_d), ptr(w_d) if w_d is not None else None, stride, ptr(self.geom),
            self.stream))
        self._keep = (arr_d, w_d)

    def _set_tti_geometry(self, metric, dualMetric):
        """TTI: [l, flat(q), (A diag(h))^-T] prepared on the device
        (_SetGeometry.py:140-142, Seismic/tti.py:113-125)."""
        if dualMetric is not None:
            raise NotImplementedError("the TTI model takes a 'metric' (no dualMetric)")
        if metric is None:
            raise ValueError(f"Missing value for key metric (model {self.model})")
        if not isinstance(metric, Metrics.Seismic.TTI):
            metric = Metrics.Seismic.TTI.from_HFM(host_array(metric))
        if metric.vdim != self.ndim:
            raise ValueError(f"TTI metric of dimension {metric.vdim} for a {self.ndim}D model")
        self.metric = metric
        lq = np.concatenate([np.asarray(host_array(metric.linear), dtype=np.float64).reshape(2, -1),
                             Metrics.flatten_symmetric_matrix(
                                 np.asarray(host_array(metric.quadratic), dtype=np.float64)).reshape(3, -1)])
        lq = lq.reshape((5,) + tuple(np.shape(host_array(metric.linear))[1:]))
        lq_d, stride = self._field_tensor(lq, (5,), 'metric (linear, quadratic)')
        A_d = None
        if metric.inverse_transformation is not None:
            A_d, astride = self._field_tensor(metric.inverse_transformation, (self.ndim, self.ndim),
                                              'metric transformation')
            if astride != stride:  # broadcast the constant one
                ncell = int(np.prod(self.shape))
                if stride == 1:
                    A_d = A_d.reshape(-1, 1).expand(A_d.numel(), ncell).contiguous()
                else:
                    lq_d = lq_d.reshape(-1, 1).expand(lq_d.numel(), ncell).contiguous()
                    stride = 1
        self.geom = empty((5 + self.ndim ** 2, self.size_tot), self.float_t)
        _lib.check(_lib.lib.agd_prepare_geometry(
            self.modelId, self.ndim, _lib.i64(self.shape), _lib.i32(self.shape_i), self.scalar,
            ctypes_double3(self.h), ptr(lq_d), ptr(A_d), stride, ptr(self.geom), self.stream))
        self._keep = (lq_d, A_d)

    def _cost_metric(self):
        """The metric with the grid steps and the cost folded in, evaluated
        at a few points (_Interface.py:202-211)."""
        metric = drift = None
        if self.model_ == 'TTI':
            metric = (self.metric.linear, self.metric.quadratic, self.metric.inverse_transformation)
        elif self.model_ == 'Diagonal':
            metric = self.metric.costs
        elif self.model_ in ('Riemann', 'Rander'):
            metric = self.metric.m
            if self.model_ == 'Rander':
                drift = self.metric.w
        return _factor.CostMetric(self.model_, self.h, self.cost_t, metric, drift)

    def _field_tensor(self, arr, core, name):
        """float32 device tensor of a field of shape core + shape (or core
        alone, broadcast), and its cell stride (1 field, 0 constant)."""
        t = device_tensor(arr, np.float32)
        if tuple(t.shape) == core:
            return t.contiguous(), 0
        if tuple(t.shape) == core + tuple(self.shape):
            return t.contiguous(), 1
        raise ValueError(f"Field {name} has incorrect dimensions. Found {tuple(t.shape)}, "
                         f"expected {core} + {self.shape}")

    # ------------------------------------------------------------ args
    def SetArgs(self):
        """_SetArgs.py:8-177: values, rhs, seeds, walls in the tiled layout."""
        if self.verbosity >= 1:
            print("Preparing the problem rhs (cost, seeds,...)")
        shp, si = _lib.i64(self.shape), _lib.i32(self.shape_i)
        esize = np.dtype(self.float_t).itemsize
        values = self.GetValue('values', default=None, help="Initial values for the eikonal solver")
        self.values_t = empty((self.size_tot,), self.float_t)
        if values is None:
            self.values_t.fill_(float('inf'))
        else:
            v = device_tensor(values, self.float_t)
            if tuple(v.shape) != tuple(self.shape):
                raise ValueError(f"Field values has incorrect dimensions {tuple(v.shape)}")
            _lib.check(_lib.lib.agd_block_expand(self.ndim, shp, si, 1, esize, ptr(v), 1,
                                                 ptr(self.values_t), float('inf'), self.stream))
        self.warmStartAllTiles = bool(self.GetValue(
            'warmStartAllTiles', default=False,
            help="Start with every tile active (use with an initial supersolution 'values')"))

        # rhs = cost, padded with +inf
        cost = device_tensor(self.cost, self.float_t)
        self.rhs_t = empty((self.size_tot,), self.float_t)
        if cost.dim() == 0:
            stride = 0
        elif tuple(cost.shape) == tuple(self.shape):
            stride = 1
        else:
            raise ValueError(f"Field cost has incorrect dimensions {tuple(cost.shape)}")
        self.cost_t = cost
        _lib.check(_lib.lib.agd_block_expand(self.ndim, shp, si, 1, esize, ptr(cost.contiguous()),
                                             stride, ptr(self.rhs_t), float('inf'), self.stream))

        nwords = (self.size_tot + 31) // 32
        self.seeds_t = zeros((nwords,), np.int32)
        self.trigger_t = zeros((nwords,), np.int32)

        seeds = self.GetValue('seeds', default=None, help="Points from where the front propagation starts")
        seedsU = self.GetValue('seeds_Unoriented', default=None,
                               help="Unoriented seeds for the curvature penalized models") \
            if self.curvature else None
        self.reverseAD = self.HasValue('sensitivity')
        self.seedIndices = None
        self.seed_coef = None
        self.seed_rows = np.zeros(0, dtype=np.int64)
        self.factor = None
        if seeds is not None:
            seeds = np.asarray(host_array(seeds), dtype=self.hfmIn.float_t)
            if seeds.ndim != 2 or seeds.shape[1] != self.ndim:
                raise ValueError(f"seeds must have shape (n, {self.ndim})")
            cont = self.hfmIn.PointFromIndex(seeds, to=True)
            seedValues = self.GetValue('seedValues', default=np.zeros(len(seeds), dtype=np.float32),
                                       help="Initial value for the front propagation")
            if ad.is_ad(seedValues):
                seedValues, self.seed_coef = seedValues.value, seedValues.coef
            else:
                svv = self.GetValue('seedValueVariation', default=None,
                                    help="First order variation of the seed values")
                if svv is not None:  # (nfwd, nseeds) -> (nseeds, nfwd)
                    self.seed_coef = np.asarray(host_array(svv)).reshape(-1, len(seeds)).T
            seedValues = np.asarray(host_array(seedValues), dtype=np.float64).reshape(-1)
            if len(seedValues) != len(seeds):
                raise ValueError("seedValues must have one entry per seed")
            seedRadius = float(self.GetValue(
                'seedRadius', default=2. if self.factoringRadius > 0 else 0.,
                help="Spread the seeds over a radius given in pixels, so as to improve accuracy."))
            if self.curvature and seedRadius > 0:
                raise NotImplementedError("seedRadius > 0 with the curvature penalized models")
            if (seedRadius > 0 or self.factoringRadius) and len(seeds) != 1:
                raise ValueError("seedRadius > 0 and factoringRadius apply to a single seed")
            cm = self._cost_metric() if (seedRadius > 0 or self.factoringRadius) else None
            if seedRadius <= 0:
                idx = np.round(cont).astype(np.int64)  # round half to even (_SetArgs.py:44-45)
                for k in range(self.ndim):
                    if self.periodic[k]:
                        idx[:, k] %= self.shape[k]
                self.seed_rows = np.arange(len(seeds))
            else:
                # _SetArgs.py:46-75: the seed becomes the cells within
                # seedRadius, valued by the mean of the cost metric norms at
                # both ends.  Each cell keeps the seed's AD row.
                idx, seedValues = _factor.spread_seed(cm, cont[0], seedRadius, self.shape,
                                                      self.periodic, seedValues[0])
                if len(idx) == 0:
                    raise ValueError("seedRadius selects no cell of the domain")
                self.seed_rows = np.zeros(len(idx), dtype=np.int64)
            if np.any(idx < 0) or np.any(idx >= np.asarray(self.shape)):
                raise ValueError("Seed outside the domain")
            if self.factoringRadius:
                a0, _ = cm.at(np.asarray(cont[:1], dtype=np.float64), self.shape, self.periodic)
                self.factor = {'radius2': self.factoringRadius ** 2,
                               'origin': np.asarray(cont[0], dtype=np.float64),
                               'metric': cm.factor_metric(a0)}
        if seedsU is not None:
            # _SetArgs.py:86-100: every orientation of each unoriented seed,
            # rounded to the nearest cell
            pU = np.asarray(host_array(seedsU), dtype=self.hfmIn.float_t).reshape(-1, 2)
            valsU = np.asarray(host_array(self.GetValue(
                'seedValues_Unoriented', default=np.zeros(len(pU)),
                help="Initial value for the front propagation")), dtype=np.float64).reshape(-1)
            if len(valsU) != len(pU):
                raise ValueError("seedValues_Unoriented must have one entry per unoriented seed")
            contU = self.hfmIn.PointFromIndex(self.hfmIn.OrientedPoints(pU).reshape(-1, 3), to=True)
            idxU = np.round(contU).astype(np.int64)
            idxU[:, 2] %= self.shape[2]
            valsU = np.tile(valsU, self.shape[2])
            if seeds is None:
                idx, seedValues = idxU, valsU
                self.seed_rows = np.arange(len(idxU))
            else:
                if self.seed_coef is not None or self.factor is not None:
                    raise NotImplementedError("seeds_Unoriented with seed AD or factorization")
                self.seed_rows = np.concatenate([self.seed_rows, len(seeds) + np.arange(len(idxU))])
                idx, seedValues = np.concatenate([idx, idxU]), np.concatenate([seedValues, valsU])
            if np.any(idx < 0) or np.any(idx >= np.asarray(self.shape)):
                raise ValueError("Seed outside the domain")
        if seeds is not None or seedsU is not None:
            self.seedIndices = idx
            self.seedValues = seedValues
            tiled = fd.tiled_index(idx, self.shape_i, self.shape_o)
            tiled_d = device_tensor(tiled, np.int64)
            vals_d = device_tensor(seedValues, np.float64)
            _lib.check(_lib.lib.agd_scatter_seeds(self.scalar, len(idx), ptr(tiled_d), ptr(vals_d),
                                                  ptr(self.rhs_t), ptr(self.seeds_t),
                                                  ptr(self.trigger_t), self.stream))
        else:
            if self.solver != 'global_iteration' and not self.warmStartAllTiles:
                raise ValueError("Missing value for key seeds")

        self.forwardAD = self.cost_coef is not None or self.seed_coef is not None
        self.values_coef = None
        self.SetChart()
        self.convergence_check = bool(self.GetValue(
            'convergence_check', default=self.curvature,
            help="After the solve, find the cells left above their scheme update and resume from them"))

        self.wallDist_t = None
        if self.walls is not None:
            walls = device_tensor(self.walls, np.uint8)
            if tuple(walls.shape) != tuple(self.shape):
                raise ValueError(f"Field walls has incorrect dimensions {tuple(walls.shape)}")
            walls_t = zeros((self.size_tot,), np.uint8)
            _lib.check(_lib.lib.agd_block_expand(self.ndim, shp, si, 1, 1, ptr(walls.contiguous()), 1,
                                                 ptr(walls_t), 0., self.stream))
            self.wallDist_t = empty((self.size_tot,), np.uint8)
            desc = self._desc()
            _lib.check(_lib.lib.agd_prepare_walls(desc, ptr(walls_t), self.wallDistBound,
                                                  ptr(self.rhs_t), ptr(self.seeds_t),
                                                  ptr(self.wallDist_t), self.stream))

        # Precomputed scheme (_SetGeometry.py:198-207): stencil weights and
        # resolved neighbour indices stored once, instead of a Selling
        # decomposition at every tile visit.  Off by default, as in the
        # reference: the stencils are rebuilt in registers.  TTI always uses
        # it (seven sub-schemes per cell exceed the register budget).
        self.scheme_w_t = self.scheme_c_t = None
        self.precompute_scheme = False
        stored = self.model_ == 'TTI' or self.curvature   # schemes always precomputed
        if self.curvature and (self.order != 1 or self.factoringRadius):
            raise NotImplementedError("the curvature models use the first-order scheme, without factorization")
        if (self.order == 1 or self.model_ == 'TTI') and (not self.factoringRadius or self.model_ == 'TTI') \
                and _lib.lib.agd_scheme_supported(
                self.modelId, self.ndim, _lib.i32(self.shape_i)):
            self.precompute_scheme = bool(self.GetValue(
                'precompute_scheme', default=stored,
                help="Precompute and store the finite difference scheme stencils"))
        if stored and not self.precompute_scheme:
            raise ValueError(f"the {self.model_} model runs from a precomputed scheme (precompute_scheme=True)")
        if self.precompute_scheme:
            # weights and codes per the library's layout (TTI and the
            # curvature models: the mix flag after the codes; TTI with order 2
            # the codes of the further neighbours x -+ 2 e_k after it)
            wc, cc = ctypes.c_int64(), ctypes.c_int64()
            _lib.check(_lib.lib.agd_scheme_lengths(self.modelId, self.ndim, self.order, ctypes.byref(wc),
                                                   ctypes.byref(cc)))
            self.scheme_w_t = empty((wc.value, self.size_tot), self.float_t)
            self.scheme_c_t = empty((cc.value, self.size_tot), np.int32)
            self.compute_scheme()

        self.nitermax_o = int(self.GetValue('nitermax_o', default=2000,
                                            help="Maximum number of iterations of the solver"))
        self.raiseOnNonConvergence = self.GetValue('raiseOnNonConvergence', default=True,
                                                   help="Raise an exception if a solver fails to converge")
        self.check_interval = int(self.GetValue(
            'check_interval', default=18,
            help="Grid passes between two convergence checks on the host"))
        self.use_graph = bool(self.GetValue(
            'hipGraph', default=True,
            help="Capture check_interval grid passes once as a hipGraph and replay it"))

    def _desc(self, niter_i=None):
        d = _lib.EikonalDesc()
        d.model = self.modelId
        d.ndim = self.ndim
        d.scalar = self.scalar
        d.solver = _SOLVER_IDS[self.solver]
        for k in range(3):
            d.shape[k] = int(self.shape[k]) if k < self.ndim else 1
            d.shape_i[k] = int(self.shape_i[k]) if k < self.ndim else 1
            d.periodic[k] = int(self.periodic[k]) if k < self.ndim else 0
            d.iso_weights[k] = float(self.iso_weights[k])
        d.niter_i = self.niter_i if niter_i is None else niter_i
        d.fim_front_width = int(self.fim_front_width or 0)
        d.atol = self.atol
        d.rtol = self.rtol
        d.nitermax_o = getattr(self, 'nitermax_o', 2000)
        d.check_interval = getattr(self, 'check_interval', 18)
        d.activate_all = int(getattr(self, 'warmStartAllTiles', False))
        d.use_graph = int(getattr(self, 'use_graph', True))
        d.time_launches = int(getattr(self, 'time_launches', False))
        d.order = int(getattr(self, 'order', 1))
        d.order2_threshold = float(getattr(self, 'order2_threshold', 0.3))
        d.curvature_relax = float(getattr(self, 'curvature_relax', 0.))
        d.convex = int(getattr(self, 'convex', False))
        if getattr(self, 'bound_active_blocks', 0):
            d.bound_active_blocks = int(self.bound_active_blocks)
            d.minchg_delta_min = float(self.minChg_delta_min)
        fac = getattr(self, 'factor', None)
        if fac is not None:
            d.factor_radius2 = fac['radius2']
            for k in range(self.ndim):
                d.factor_origin[k] = float(fac['origin'][k])
            for k, v in enumerate(fac['metric']):
                d.factor_metric[k] = float(v)
        return d

    def compute_scheme(self):
        """Fill the precomputed scheme from the tiled geometry and walls."""
        _lib.check(_lib.lib.agd_eikonal_scheme(self._desc(), self._io(), self.stream))

    def _io(self):
        io = _lib.EikonalIO()
        io.values = ptr(self.values_t)
        io.geom = ptr(self.geom) if self.geom is not None else None
        io.rhs = ptr(self.rhs_t)
        io.seeds = ptr(self.seeds_t)
        io.trigger = ptr(self.trigger_t)
        io.wall_dist = ptr(self.wallDist_t) if self.wallDist_t is not None else None
        sw, sc = getattr(self, 'scheme_w_t', None), getattr(self, 'scheme_c_t', None)
        io.scheme_weights = ptr(sw)
        io.scheme_codes = ptr(sc)
        io.scheme_weights_len = sw.numel() if sw is not None else 0
        io.scheme_codes_len = sc.numel() if sc is not None else 0
        return io

    def solver_handle(self):
        """The solver handle (worklist workspace, cached pass graph) of this
        run, bound to its stream; created on first use."""
        h = getattr(self, '_solver_handle', None)
        if h is None:
            h = self._solver_handle = _lib.Solver(self.stream, self._desc())
        return h

    # ------------------------------------------------------------ charts
    def SetChart(self):
        """_StopCriterion.py:93-165: the mapping between local charts, from
        physical coordinates to continuous grid indices of the last ndim_s
        axes (broadcast over the first ones)."""
        self.hasChart = self.HasValue('chart_mapping')
        if not self.hasChart:
            return
        mapping = self.GetValue('chart_mapping', default=None,
                                help="Mapping from one local chart to another, for eikonal equations defined "
                                     "on manifolds. (Please set to NaN inside the walls.)")
        self.chart_nitermax = int(self.GetValue(
            'chart_nitermax', default=5,
            help="Number of times the boundary conditions are updated in the eikonal solver, "
                 "with values from other local charts."))
        m = np.asarray(host_array(mapping), dtype=np.float64)
        ndim_s = m.ndim - 1
        ndim_b = self.ndim - ndim_s
        if ndim_s < 1 or ndim_b < 0 or tuple(m.shape[1:]) != tuple(self.shape[ndim_b:]) or len(m) != ndim_s:
            raise ValueError(f"Inconsistent shape of field chart_mapping : {m.shape}")
        if self.forwardAD or self.reverseAD:
            raise ValueError("Sorry, forward and reversed AD not yet supported with charts")
        lo, _ = self.hfmIn.corners
        lo = np.asarray(host_array(lo), dtype=np.float64)[ndim_b:]
        h = np.asarray(self.h, dtype=np.float64)[ndim_b:]
        shp = (ndim_s,) + (1,) * ndim_s
        m = (m - lo.reshape(shp)) / h.reshape(shp) - 0.5
        self.chart_ndim_s = ndim_s
        self.chart_mapping_t = device_tensor(np.ascontiguousarray(m), self.float_t)
        # A paste lowers a value only by more than the solve's own tolerance:
        # smaller decreases are the rounding of the interpolation, and each
        # would cost a resumed solve (the reference pastes any decrease).
        self.chart_atol, self.chart_rtol = (float(t) for t in self.GetValue(
            'chart_tolerance', default=(self.atol, self.rtol),
            help="(atol, rtol): decreases a paste between charts must exceed (0, 0: any decrease)"))

    def _chart_paste(self, scratch, trig, cnt):
        """One paste of the values between charts; the number of cells lowered."""
        trig.zero_()
        cnt.zero_()
        _lib.check(_lib.lib.agd_chart_paste(self._desc(), ptr(self.values_t), ptr(self.chart_mapping_t),
                                            self.chart_ndim_s, self.chart_atol, self.chart_rtol, ptr(scratch),
                                            ptr(trig), ptr(cnt), self.stream))
        return int(cnt.item())

    # ------------------------------------------------------------ solve
    def Solve(self):
        """_solvers.py:15-57; with charts, after each converged solve the
        values are pasted between charts and the solve continues from the
        lowered cells, up to chart_nitermax times (_StopCriterion.py:60-80)."""
        if self.verbosity >= 1:
            print("Running the eikonal GPU kernel")
        desc, io, st = self._desc(), self._io(), _lib.EikonalStats()
        torch.cuda.synchronize()
        t0 = time.time()
        status = self.solver_handle().solve(desc, io, st)
        totals = st.as_dict()

        def resume(trig):
            # continue the solve from the tiles of the marked cells
            d2, io2, st2 = self._desc(), self._io(), _lib.EikonalStats()
            d2.activate_all = 0
            io2.trigger = ptr(trig)
            s = self.solver_handle().solve(d2, io2, st2)
            for k in ('niter_o', 'tile_updates', 'cell_updates', 'launches', 'gpu_ms', 'kernel_ms'):
                totals[k] += getattr(st2, k)
            totals['converged'] = st2.converged
            return s

        # The curvature models' stencils reach two tiles away, further than
        # the face neighbours a changed tile re-lists: cells left above
        # their update are found by a residual pass and the solve resumes
        # from them (no counterpart in the reference).
        if self.convergence_check and status == _lib.AGD_OK:
            trig = zeros(((self.size_tot + 31) // 32,), np.int32)
            rounds = 0
            while rounds < 8:
                trig.zero_()
                nstale = self.solver_handle().check_stale(self._desc(), self._io(), self.atol, self.rtol, ptr(trig))
                if nstale == 0:
                    break
                rounds += 1
                status = resume(trig)
                if status != _lib.AGD_OK:
                    break
            totals['convergence_check_rounds'] = rounds
            del trig
        if getattr(self, 'hasChart', False):
            scratch = empty((self.size_tot,), self.float_t)
            trig = zeros(((self.size_tot + 31) // 32,), np.int32)
            cnt = zeros((1,), np.int32)
            npaste = 0
            while status == _lib.AGD_OK and npaste < self.chart_nitermax:
                npaste += 1
                if self._chart_paste(scratch, trig, cnt) == 0:
                    break
                status = resume(trig)
            totals['chart_pastes'] = npaste
            del scratch, trig, cnt
        wall = time.time() - t0
        self.stats['eikonal'] = {**totals, 'time': wall, 'gpu_time': totals['gpu_ms'] * 1e-3}
        if self.verbosity >= 1:
            print(f"GPU kernel eikonal ran for {wall} seconds, and {st.niter_o} iterations.")
        if status == _lib.AGD_NOCONV:
            msg = (f"Solver {self.solver} for kernel eikonal did not reach convergence after "
                   f"maximum allowed number {self.nitermax_o} of iterations")
            if self.raiseOnNonConvergence:
                raise ValueError(msg)
            self.Warn(msg)
        else:
            _lib.check(status)
        self._values = None

    @property
    def values(self):
        if self._values is None:
            out = empty(tuple(self.shape), self.float_t)
            _lib.check(_lib.lib.agd_block_squeeze(self.ndim, _lib.i64(self.shape), _lib.i32(self.shape_i),
                                                  1, np.dtype(self.float_t).itemsize,
                                                  ptr(self.values_t), ptr(out), self.stream))
            if self.export_t != self.float_t:
                from ._device import tdtype
                out = out.to(tdtype(self.export_t))
            self._values = out
        if self.values_coef is not None:
            return ad.denseAD(self._output(self._values), self._output(self.values_coef))
        return self._output(self._values)

    def _output(self, t):
        return t.cpu().numpy() if self.host_outputs else t

    def _squeeze(self, t_tiled, ncomp, dtype):
        out = empty((ncomp,) + tuple(self.shape), dtype)
        _lib.check(_lib.lib.agd_block_squeeze(self.ndim, _lib.i64(self.shape), _lib.i32(self.shape_i),
                                              ncomp, np.dtype(dtype).itemsize, ptr(t_tiled),
                                              ptr(out), self.stream))
        return out

    # ------------------------------------------------------------ flow
    def PostProcess(self):
        """_PostProcess.py:46-95: one global pass exporting the upwind flow.
        Geodesics alone need no exported flow when it is recomputed online
        (_Kernel.py:148-150, 223-224)."""
        self.geodesic_online_flow = bool(self.GetValue(
            'geodesic_online_flow', default=False,
            help="Compute the flow online when extracting geodesics (saves memory)")) if self.hasTips else False
        self.flow_needed = (self.exportGeodesicFlow or self.reverseAD or self.forwardAD
                            or (self.hasTips and not self.geodesic_online_flow)
                            or getattr(self, 'keep_flow', False))
        if not self.flow_needed:
            return
        if self.verbosity >= 1:
            print("Post-Processing")
        nsym = scheme_slots(self.model_, self.ndim)
        self.nsym = nsym
        self.flow_weights_t = empty((nsym, self.size_tot), self.float_t)
        self.flow_weightsum_t = empty((self.size_tot,), self.float_t)
        self.flow_indices_t = empty((nsym, self.size_tot), np.int32)
        self.flow_vector_t = empty((self.ndim, self.size_tot), self.float_t)
        f = _lib.FlowIO()
        f.flow_weights = ptr(self.flow_weights_t)
        f.flow_weightsum = ptr(self.flow_weightsum_t)
        f.flow_indices = ptr(self.flow_indices_t)
        f.flow_vector = ptr(self.flow_vector_t)
        t0 = time.time()
        _lib.check(_lib.lib.agd_eikonal_flow(self._desc(niter_i=1), self._io(), f, self.stream))
        torch.cuda.synchronize()
        self.stats['flow'] = {'niter_o': 1, 'time': time.time() - t0}
        self._values = None
        # _PostProcess.py:80-91: with a drift, the flow is divided by its
        # metric norm (the flow on the adimensionized grid), which the AD
        # right-hand sides and sensitivities also carry.
        self.flow_normalization = None
        if self.exportGeodesicFlow or self.model_ == 'Rander':
            fv = self._squeeze(self.flow_vector_t, self.ndim, self.float_t)
            if self.model_ == 'Rander':
                self.flow_normalization = self._rander_flow_normalization(fv)
                fv = fv / self.flow_normalization
            if self.exportGeodesicFlow:
                h = torch.as_tensor(self.h, dtype=fv.dtype, device=fv.device).reshape((self.ndim,) + (1,) * self.ndim)
                self.hfmOut['flow'] = self._output(fv * (-h))

    def _seed_mask(self):
        """Cell-major bool mask of the seed bits (seeds and walls)."""
        words = self.seeds_t.to(torch.int64) & 0xffffffff
        bits = ((words.reshape(-1, 1) >> torch.arange(32, device=words.device)) & 1).to(torch.uint8)
        bits = bits.reshape(-1)[:self.size_tot].contiguous()
        return self._squeeze(bits, 1, np.uint8)[0].bool()

    def _rander_flow_normalization(self, fv):
        """metric.norm(-flow) with the metric in grid units (with_costs(h):
        N_h(v) = sqrt(<h v, m h v>) + <w, h v>), 1 on the seeds and walls
        (_PostProcess.py:86-87, Metrics/rander.py:30-32)."""
        arr_d, w_d = self._keep
        d, n = self.ndim, int(np.prod(self.shape))
        m = arr_d.to(fv.dtype).reshape(d, d, -1)
        w = w_d.to(fv.dtype).reshape(d, -1)
        h = torch.as_tensor(self.h, dtype=fv.dtype, device=fv.device)
        v = -fv.reshape(d, n) * h[:, None]
        mv = (m * v[None, :, :]).sum(1)
        norm = torch.sqrt((v * mv).sum(0)) + (w * v).sum(0)
        norm = torch.where(self._seed_mask().reshape(-1), torch.ones_like(norm), norm)
        return norm.reshape(tuple(self.shape))

    # ------------------------------------------------------------ geodesics
    _GEODESIC_STOP = ('Continue', 'AtSeed', 'InWall', 'Stationnary', 'PastSeed', 'VanishingFlow')

    def GetGeodesics(self):
        """_GetGeodesics.py:9-147: backtrack the minimal geodesics from the
        tips along the upwind flow (agd_geodesic_ode), in chunks of
        geodesic_typical_length points per launch until every geodesic has
        stopped or geodesic_max_length points are reached."""
        if not self.hasTips:
            return
        if self.verbosity >= 1:
            print("Computing the geodesics")
        tips = np.zeros((0, self.ndim), self.float_t) if self.tips is None else \
            self.hfmIn.PointFromIndex(self.tips, to=True).astype(self.float_t)
        if self.tips_Unoriented is not None:
            # _GetGeodesics.py:13-25: each unoriented tip starts from the
            # cell of its most favourable orientation
            pts = self.hfmIn.OrientedPoints(self.tips_Unoriented)      # (ntheta, n, 3)
            idxU = np.round(self.hfmIn.PointFromIndex(pts, to=True)).astype(np.int64)
            idxU[..., 2] %= self.shape[2]
            vals = self.values
            vals = vals.cpu().numpy() if isinstance(vals, torch.Tensor) else np.asarray(ad.remove_ad(vals))
            vU = vals[tuple(np.moveaxis(idxU, -1, 0))]                   # (ntheta, n)
            best = np.argmin(vU, axis=0)
            tipsU = idxU[best, np.arange(idxU.shape[1])].astype(self.float_t)
            tips = np.concatenate([tips, tipsU])
        step = float(self.GetValue('geodesic_step', default=0.25,
                                   help='Step size, in pixels, for the geodesic ODE solver'))
        eucl_delay = int(self.GetValue(
            'geodesic_PastSeed_delay', default=np.sqrt(self.ndim) / step,
            help="Delay, in iterations, for the 'PastSeed' stopping criterion of the geodesic ODE solver"))
        nymin_delay = int(self.GetValue(
            'geodesic_Stationnary_delay', default=8. * np.sqrt(self.ndim) / step,
            help="Delay, in iterations, for the 'Stationnary' stopping criterion of the geodesic ODE solver"))
        shape_max = int(np.max(self.shape))
        max_len = int(max(40, 20 * shape_max / step))
        max_len = self.GetValue('geodesic_max_length', default=max_len,
                                help="Maximum allowed length of geodesics.")
        typical_len = int(max(40, 0.5 * shape_max / step))
        typical_len = int(self.GetValue('geodesic_typical_length', default=min(max_len, typical_len),
                                        help="Typical expected length of geodesics (number of points)."))
        causality = float(self.GetValue('geodesic_causalityTolerance', default=4.,
                                        help="Used in criterion for rejecting points in flow interpolation"))
        eucl_bound = self.GetValue('geodesic_targetTolerance', default=12 if self.curvature else 6,
                                   help="Tolerance, in pixels, for declaring a seed as reached.")
        # Distance to the seeds, in fifths of a pixel, up to eucl_bound pixels
        # (the trigger bits: seeds without the walls).
        desc = self._desc()
        eucl = empty((self.size_tot,), np.uint8)
        tmp = empty((self.size_tot,), np.uint8)
        mult = 5
        _lib.check(_lib.lib.agd_geodesic_eucl(desc, ptr(self.trigger_t), int(np.ceil(eucl_bound)), mult,
                                              int(mult * eucl_bound), ptr(eucl), ptr(tmp), self.stream))
        del tmp
        chart_var = float('inf')
        if getattr(self, 'hasChart', False):
            # _GetGeodesics.py:66-71, _Kernel.py:189-201: the cells where
            # geodesics jump to another chart carry the eucl code 254
            jump = self.GetValue('chart_jump', help="Where the geodesics should jump "
                                 "to another local chart of the manifold")
            jump = np.ascontiguousarray(np.broadcast_to(np.asarray(host_array(jump), dtype=np.uint8),
                                                        tuple(self.shape)))
            jump_c = device_tensor(jump, np.uint8)
            jump_t = zeros((self.size_tot,), np.uint8)
            _lib.check(_lib.lib.agd_block_expand(self.ndim, _lib.i64(self.shape), _lib.i32(self.shape_i), 1, 1,
                                                 ptr(jump_c), 1, ptr(jump_t), 0., self.stream))
            eucl.masked_fill_(eucl == _lib.EUCL_CHART, 255)
            eucl.masked_fill_(jump_t != 0, _lib.EUCL_CHART)
            dev = self.GetValue('chart_jump_deviation', default=np.inf,
                                help="Do not interpolate the jump coordinates, among pixel corners, "
                                     " if their (adimensionized) standard deviation exceeds this threshold. "
                                     "(Use if chart_mapping is discontinuous. Typical value : 5.) ")
            dev = 5. if dev is True else float(dev)
            chart_var = dev * dev
        pastseed_delay = int(self.GetValue(
            'pastseed_delay', default=0,
            help="Number of geodesic backtracking steps before checking the PastSeed criterion "
                 "(use if seed and tip are close)"))
        self.GetValue('geodesic_block_size', default=64,
                      help="Block size for the GPU based geodesic solver (one wave64 here)")

        ngeo = len(tips)
        stopping = ["Stopping criterion"] * ngeo
        geodesics = [[tip.reshape(1, -1)] for tip in tips]
        corresp = list(range(ngeo))
        geo_max_it = int(np.ceil(max_len / typical_len)) if np.isfinite(max_len) else None
        it = 0
        dev = self.values_t.device
        t0 = time.time()
        while corresp:
            if geo_max_it is not None and it >= geo_max_it:
                self.Warn("Geodesic solver failed to converge, or geodesic has too many points"
                          " (in latter case, try setting 'geodesic_max_length':np.inf)")
                break
            it += 1
            n = len(corresp)
            starts = np.stack([geodesics[i][-1][-1, :] for i in corresp]).astype(self.float_t)
            x_s = torch.full((n, typical_len, self.ndim), float('nan'), dtype=self.values_t.dtype, device=dev)
            x_s[:, 0, :] = torch.as_tensor(starts, device=dev)
            len_s = torch.full((n,), -1, dtype=torch.int32, device=dev)
            stop_s = torch.full((n,), -1, dtype=torch.int8, device=dev)
            gd = _lib.GeodesicDesc()
            gd.ngeo, gd.max_len = n, typical_len
            gd.eucl_delay, gd.nymin_delay, gd.pastseed_delay = eucl_delay, nymin_delay, pastseed_delay
            gd.step, gd.causality_tolerance = step, causality
            if getattr(self, 'hasChart', False):
                gd.chart_ndim_s = self.chart_ndim_s
                gd.chart_mapping = ptr(self.chart_mapping_t)
                gd.chart_jump_variance = chart_var
            else:
                gd.chart_jump_variance = float('inf')
            if self.geodesic_online_flow:
                # the flow of each cell the ODE enters, recomputed from the scheme
                _lib.check(_lib.lib.agd_geodesic_ode_online(desc, self._io(), gd, ptr(eucl), ptr(x_s),
                                                            ptr(len_s), ptr(stop_s), self.stream))
            else:
                _lib.check(_lib.lib.agd_geodesic_ode(desc, gd, ptr(self.values_t), ptr(self.flow_vector_t),
                                                     ptr(self.flow_weightsum_t), ptr(eucl), ptr(x_s),
                                                     ptr(len_s), ptr(stop_s), self.stream))
            x_h, len_h, stop_h = x_s.cpu().numpy(), len_s.cpu().numpy(), stop_s.cpu().numpy()
            nxt = []
            for i, x, ln, sp in zip(corresp, x_h, len_h, stop_h):
                geodesics[i].append(x[1:int(ln)])
                if sp != 0:
                    stopping[i] = self._GEODESIC_STOP[int(sp)]
                else:
                    nxt.append(i)
            corresp = nxt
            pastseed_delay = 0
        self.stats['geodesic'] = {'time': time.time() - t0, 'launches': it}
        geos = [self.hfmIn.PointFromIndex(np.concatenate(g, axis=0)).T for g in geodesics]
        ntips = 0 if self.tips is None else len(self.tips)
        if self.tips is not None:
            self.hfmOut['geodesics'] = geos[:ntips]
        if self.tips_Unoriented is not None:
            self.hfmOut['geodesics_Unoriented'] = geos[ntips:]
        self.hfmOut['geodesic_stopping_criteria'] = stopping

    # ------------------------------------------------------------ AD
    def SolveAD(self):
        """_PostProcess.py:155-219: forward then reverse differentiation."""
        if not (self.forwardAD or self.reverseAD):
            return
        if self.order == 2 or self.factoringRadius:
            # _PostProcess.py:165-166: the linearised scheme is the flow's.
            self.Warn("Eikonal AD ignores order2 and source factorization")
        if self.forwardAD:
            self.values_coef = self.forward_variation()
        if self.reverseAD:
            sens = self.GetValue('sensitivity', help='Reverse automatic differentiation')
            lam, seedSens = self.reverse(sens)
            self.costSensitivity = lam
            self.hfmOut['costSensitivity'] = self._output(lam)
            # _PostProcess.py:217-219: each seed collects the sensitivities of
            # the cells it was spread to (one cell per seed without spreading).
            self.hfmOut['seedValueSensitivity'] = self._output(seedSens)

    def _expand_components(self, x_c, pad):
        """(n, *shape) cell-major device tensor -> (n, size_tot) tiled."""
        n = int(x_c.shape[0])
        out = empty((n, self.size_tot), self.float_t)
        _lib.check(_lib.lib.agd_block_expand(self.ndim, _lib.i64(self.shape), _lib.i32(self.shape_i),
                                             n, np.dtype(self.float_t).itemsize, ptr(x_c), 1,
                                             ptr(out), pad, self.stream))
        return out

    def _linear_tolerances(self, rhs_c):
        """linear_rtol / linear_atol, set by the first linear solve
        (_PostProcess.py:126-138)."""
        if not hasattr(self, 'linear_rtol'):
            res = np.finfo(self.float_t).resolution
            self.linear_rtol = float(self.GetValue('linear_rtol', default=float(res * 5),
                                                   help="Relative convergence tolerance for the linear systems"))
        if not hasattr(self, 'linear_atol'):
            linear_atol = self.GetValue('linear_atol', default=None,
                                        help="Absolute convergence tolerance for the linear systems")
            if linear_atol is None:
                fin = torch.isfinite(rhs_c)
                mean = float(rhs_c[fin].abs().double().mean().item()) if bool(fin.any()) else 0.
                linear_atol = self.linear_rtol * mean
                self.hfmOut['keys']['default']['linear_atol'] = linear_atol
            self.linear_atol = float(linear_atol)

    def _linear_solve(self, entry, name, rhs_c):
        """Tile Gauss-Seidel solve of the linearised scheme for the
        right-hand sides rhs_c (n, *shape); returns (*shape, n)."""
        nrhs = int(rhs_c.shape[0])
        rhs_t = self._expand_components(rhs_c, float('nan'))
        self._linear_tolerances(rhs_c)
        sol_t = empty((nrhs, self.size_tot), self.float_t)
        st = _lib.EikonalStats()
        status = entry(self._desc(), ptr(self.values_t), ptr(self.flow_weights_t),
                       ptr(self.flow_indices_t), ptr(self.flow_weightsum_t), nrhs, ptr(rhs_t),
                       ptr(sol_t), self.linear_atol, self.linear_rtol, st, self.stream)
        self.stats[name] = {**st.as_dict(), 'time': st.gpu_ms * 1e-3}
        if status == _lib.AGD_NOCONV:
            msg = f"Solver for kernel {name} did not reach convergence"
            if self.raiseOnNonConvergence:
                raise ValueError(msg)
            self.Warn(msg)
        else:
            _lib.check(status)
        return torch.movedim(self._squeeze(sol_t, nrhs, self.float_t), 0, -1).contiguous()

    def _seed_pos(self, dev):
        return tuple(torch.as_tensor(self.seedIndices[:, k], device=dev) for k in range(self.ndim))

    def forward_variation(self):
        """First-order variation of the values (_PostProcess.py:181-191):
        solves (D - W) du = r with r = cost * costVariation off the seeds,
        the seed value variation on the seeds and 0 in the walls."""
        cc = None if self.cost_coef is None else device_tensor(self.cost_coef, self.float_t)
        sc = None if self.seed_coef is None else device_tensor(self.seed_coef, self.float_t)
        nf = int((cc if cc is not None else sc).shape[-1])
        if cc is not None and sc is not None and int(sc.shape[-1]) != nf:
            raise ValueError("costVariation and seedValueVariation have different AD sizes")
        if cc is not None:
            if tuple(cc.shape) != tuple(self.shape) + (nf,):
                raise ValueError(f"costVariation shape {tuple(cc.shape)} must be {self.shape} + (n,)")
            rhs = cc * self.cost_t.reshape(self.cost_t.shape + (1,))
        else:
            rhs = zeros(tuple(self.shape) + (nf,), self.float_t)
        if self.seedIndices is not None:
            rows = torch.as_tensor(self.seed_rows, device=rhs.device)
            rhs[self._seed_pos(rhs.device)] = sc[rows] if sc is not None else 0.
        if self.walls is not None:
            rhs[device_tensor(self.walls, np.bool_)] = 0.
        if getattr(self, "flow_normalization", None) is not None:      # _PostProcess.py:185
            rhs = rhs * self.flow_normalization.reshape(self.flow_normalization.shape + (1,))
        rhs_c = torch.movedim(rhs, -1, 0).contiguous()
        return self._linear_solve(_lib.lib.agd_eikonal_forward_ad, 'forwardAD', rhs_c)

    def reverse(self, sens):
        """Reverse sensitivities for sens of shape (*shape, nrhs)
        (_PostProcess.py:194-219): returns (costSensitivity, with the seeds
        zeroed, and the seeds' rows)."""
        sens = device_tensor(sens, self.float_t)
        if tuple(sens.shape[:-1]) != tuple(self.shape):
            raise ValueError(f"Reverse AD rhs shape {tuple(sens.shape)} does not start with {self.shape}")
        sens_c = torch.movedim(sens, -1, 0).contiguous()
        lam = self._linear_solve(_lib.lib.agd_eikonal_adjoint, 'reverseAD', sens_c)
        if self.seedIndices is None:
            return lam, lam.new_zeros((0, int(sens.shape[-1])))
        pos = self._seed_pos(lam.device)
        cellSens = lam[pos].clone()
        lam[pos] = 0
        if getattr(self, "flow_normalization", None) is not None:      # _PostProcess.py:213-214
            lam = lam / self.flow_normalization.reshape(self.flow_normalization.shape + (1,))
        nseeds = int(self.seed_rows.max()) + 1 if len(self.seed_rows) else 0
        seedSens = cellSens.new_zeros((nseeds, int(cellSens.shape[-1])))
        seedSens.index_add_(0, torch.as_tensor(self.seed_rows, device=lam.device), cellSens)
        return lam, seedSens

    # ------------------------------------------------------------ outputs
    def FinalCheck(self):
        """_Interface.py:162-174."""
        if self.GetValue('exportValues', False, help="Return the solution numerical values"):
            self.hfmOut['values'] = self.values
        self.extractValues = self.GetValue('extractValues', False,
                                           help="Return the solution numerical values separately from other data")
        self.retself = self.GetValue('retself', False, help="Return the class instance that did the work")
        self.hfmOut['stats'] = self.stats
        self.hfmOut['solverGPUTime'] = self.stats['eikonal']['time']
        self.hfmOut['keys']['unused'] = list(set(self.hfmIn.keys()) - set(self.hfmOut['keys']['used']))
        if self.verbosity >= 1 and self.hfmOut['keys']['unused']:
            print(f"!! Warning !! Unused keys from user : {self.hfmOut['keys']['unused']}")


def ctypes_double3(h):
    import ctypes
    arr = (ctypes.c_double * 3)(*([float(x) for x in h] + [1.] * (3 - len(h))))
    return arr
