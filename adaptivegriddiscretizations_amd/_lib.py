"""ctypes binding of libagd_hip.so, the C ABI declared in include/agd_hip.h.

The library is the only compute path of this package: if it cannot be
loaded, importing this module raises (there is no CPU fallback).
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_void_p

from . import _build

c_int32_p = POINTER(c_int32)
c_int64_p = POINTER(c_int64)
c_double_p = POINTER(c_double)
c_uint8_p = POINTER(ctypes.c_uint8)
c_uint32_p = POINTER(ctypes.c_uint32)

# agd_status
AGD_OK, AGD_EINVAL, AGD_EHIP, AGD_NOCONV, AGD_ENOMEM = 0, -1, -2, -3, -4
# agd_model
ISOTROPIC, DIAGONAL, RIEMANN, RANDER, TTI = 0, 1, 2, 3, 4
REEDSSHEPP2, REEDSSHEPPFWD2, ELASTICA2, DUBINS2 = 5, 6, 7, 8
EUCL_CHART = 254
# agd_scalar
F32, F64 = 0, 1
# agd_solver_kind
AGSI, FIM, GLOBAL = 0, 1, 2


class EikonalDesc(ctypes.Structure):
    _fields_ = [
        ("model", c_int32), ("ndim", c_int32), ("scalar", c_int32), ("solver", c_int32),
        ("shape", c_int64 * 3), ("shape_i", c_int32 * 3), ("periodic", c_int32 * 3),
        ("niter_i", c_int32), ("fim_front_width", c_int32),
        ("atol", c_double), ("rtol", c_double), ("nitermax_o", c_int64),
        ("iso_weights", c_double * 3), ("check_interval", c_int32), ("activate_all", c_int32),
        ("time_launches", c_int32), ("use_graph", c_int32),
        ("order", c_int32), ("convex", c_int32), ("order2_threshold", c_double),
        ("factor_radius2", c_double), ("factor_origin", c_double * 3), ("factor_metric", c_double * 14),
        ("bound_active_blocks", c_int64), ("minchg_delta_min", c_double),
        ("curvature_relax", c_double),
    ]


class EikonalIO(ctypes.Structure):
    _fields_ = [
        ("values", c_void_p), ("geom", c_void_p), ("rhs", c_void_p),
        ("seeds", c_void_p), ("trigger", c_void_p), ("wall_dist", c_void_p),
        ("scheme_weights", c_void_p), ("scheme_codes", c_void_p),
        ("scheme_weights_len", c_int64), ("scheme_codes_len", c_int64),
    ]


class BatchInstance(ctypes.Structure):
    _fields_ = [("shape", c_int64 * 3), ("iso_weights", c_double * 3), ("atol", c_double),
                ("rtol", c_double)]


class EikonalStats(ctypes.Structure):
    _fields_ = [
        ("niter_o", c_int64), ("tile_updates", c_int64), ("cell_updates", c_int64),
        ("launches", c_int64), ("gpu_ms", c_double), ("converged", c_int32),
        ("grid_blocks", c_int32), ("kernel_ms", c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class GeodesicDesc(ctypes.Structure):
    _fields_ = [
        ("ngeo", c_int32), ("max_len", c_int32), ("eucl_delay", c_int32), ("nymin_delay", c_int32),
        ("pastseed_delay", c_int32), ("chart_ndim_s", c_int32), ("step", c_double),
        ("causality_tolerance", c_double), ("chart_mapping", c_void_p), ("chart_jump_variance", c_double),
    ]


class FlowIO(ctypes.Structure):
    _fields_ = [
        ("flow_weights", c_void_p), ("flow_weightsum", c_void_p),
        ("flow_indices", c_void_p), ("flow_vector", c_void_p),
    ]


# name -> (restype, argtypes).  Kept in the order of include/agd_hip.h.
SIGNATURES = {
    "agd_last_error": (c_char_p, []),
    "agd_version": (c_int32, []),
    "agd_tile_supported": (c_int32, [c_int32, c_int32, c_int32_p]),
    "agd_batch_tile_supported": (c_int32, [c_int32, c_int32, c_int32_p]),
    "agd_device_count": (c_int32, [c_int32_p]),
    "agd_solver_create": (c_int32, [POINTER(EikonalDesc), c_void_p, POINTER(c_void_p)]),
    "agd_solver_destroy": (c_int32, [c_void_p]),
    "agd_solver_solve": (c_int32, [c_void_p, POINTER(EikonalDesc), POINTER(EikonalIO),
                                   POINTER(EikonalStats)]),
    "agd_solver_solve_batched": (c_int32, [c_void_p, POINTER(EikonalDesc), c_int32,
                                           POINTER(BatchInstance), POINTER(EikonalIO),
                                           POINTER(EikonalStats)]),
    "agd_solver_check": (c_int32, [c_void_p, POINTER(EikonalDesc), POINTER(EikonalIO), c_double,
                                   c_double, POINTER(c_int64)]),
    "agd_solver_check_stale": (c_int32, [c_void_p, POINTER(EikonalDesc), POINTER(EikonalIO), c_double,
                                         c_double, c_void_p, POINTER(c_int64)]),
    "agd_solver_repair": (c_int32, [c_void_p, POINTER(EikonalDesc), POINTER(EikonalIO), c_double,
                                    c_double, c_int32, POINTER(c_int64), POINTER(c_int32)]),
    "agd_eikonal_solve": (c_int32, [POINTER(EikonalDesc), POINTER(EikonalIO),
                                    POINTER(EikonalStats), c_void_p]),
    "agd_eikonal_solve_noalloc": (c_int32, [POINTER(EikonalDesc), POINTER(EikonalIO),
                                            POINTER(EikonalStats), c_void_p]),
    "agd_eikonal_flow": (c_int32, [POINTER(EikonalDesc), POINTER(EikonalIO), POINTER(FlowIO),
                                   c_void_p]),
    "agd_eikonal_solve_batched": (c_int32, [POINTER(EikonalDesc), c_int32, POINTER(BatchInstance),
                                            POINTER(EikonalIO), POINTER(EikonalStats), c_void_p]),
    "agd_eikonal_check": (c_int32, [POINTER(EikonalDesc), POINTER(EikonalIO), c_double, c_double,
                                    POINTER(c_int64), c_void_p]),
    "agd_scheme_supported": (c_int32, [c_int32, c_int32, c_int32_p]),
    "agd_scheme_lengths": (c_int32, [c_int32, c_int32, c_int32, c_int64_p, c_int64_p]),
    "agd_eikonal_scheme": (c_int32, [POINTER(EikonalDesc), POINTER(EikonalIO), c_void_p]),
    "agd_eikonal_adjoint": (c_int32, [POINTER(EikonalDesc), c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_int32, c_void_p, c_void_p, c_double,
                                      c_double, POINTER(EikonalStats), c_void_p]),
    "agd_eikonal_forward_ad": (c_int32, [POINTER(EikonalDesc), c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_int32, c_void_p, c_void_p, c_double,
                                         c_double, POINTER(EikonalStats), c_void_p]),
    "agd_geodesic_eucl": (c_int32, [POINTER(EikonalDesc), c_void_p, c_int32, c_int32, c_int32,
                                    c_void_p, c_void_p, c_void_p]),
    "agd_geodesic_ode": (c_int32, [POINTER(EikonalDesc), POINTER(GeodesicDesc), c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "agd_geodesic_ode_online": (c_int32, [POINTER(EikonalDesc), POINTER(EikonalIO), POINTER(GeodesicDesc),
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "agd_chart_paste": (c_int32, [POINTER(EikonalDesc), c_void_p, c_void_p, c_int32, c_double, c_double,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "agd_block_expand": (c_int32, [c_int32, c_int64_p, c_int32_p, c_int32, c_int32, c_void_p,
                                   c_int64, c_void_p, c_double, c_void_p]),
    "agd_block_squeeze": (c_int32, [c_int32, c_int64_p, c_int32_p, c_int32, c_int32, c_void_p,
                                    c_void_p, c_void_p]),
    "agd_prepare_geometry": (c_int32, [c_int32, c_int32, c_int64_p, c_int32_p, c_int32,
                                       c_double_p, c_void_p, c_void_p, c_int64, c_void_p,
                                       c_void_p]),
    "agd_scatter_seeds": (c_int32, [c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    "agd_prepare_walls": (c_int32, [POINTER(EikonalDesc), c_void_p, c_int32, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "agd_stencil_export": (c_int32, [c_int32, c_int32, c_int32, c_int64, c_void_p, c_double_p,
                                     c_void_p, c_void_p, c_void_p]),
    "agd_nn_conv3x3": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                                 c_int32, c_int32, c_void_p]),
    "agd_nn_conv3x3_wgrad": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                       c_int32, c_int32, c_void_p]),
    "agd_nn_conv3x3_flip": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_void_p]),
    "agd_nn_relu_mask": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p]),
    "agd_nn_pool2": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32,
                               c_void_p]),
    "agd_nn_up2": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32,
                             c_void_p]),
    "agd_nn_add": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "agd_nn_mse": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_double,
                             c_double, c_void_p, c_void_p]),
    "agd_nn_adam": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_double,
                              c_double, c_double, c_double, c_int32, c_void_p]),
    "agd_nn_warm_values": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_int32, c_double,
                                     c_void_p, c_void_p, c_int32, c_int32, c_void_p]),
    "agd_malloc": (c_int32, [POINTER(c_void_p), c_int64]),
    "agd_free": (c_int32, [c_void_p]),
    "agd_memcpy_htod": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p]),
    "agd_memcpy_dtoh": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p]),
    "agd_memset": (c_int32, [c_void_p, c_int32, c_int64, c_void_p]),
    "agd_stream_synchronize": (c_int32, [c_void_p]),
}


class AGDError(RuntimeError):
    """A call into libagd_hip.so failed."""

    def __init__(self, status, message):
        super().__init__(f"libagd_hip status {status}: {message}")
        self.status = status


class NonConvergence(AGDError):
    pass


def hip_runtimes_loaded():
    """Paths of the HIP runtime libraries mapped in this process."""
    with open("/proc/self/maps") as f:
        return sorted({line.split()[-1] for line in f if "libamdhip64" in line})


def _load():
    path = os.environ.get("AGD_HIP_LIB", _build.LIB_PATH)
    if not os.path.exists(path):
        raise ImportError(
            f"libagd_hip.so not found at {path}: build it first "
            "(python -c 'import __graft_entry__; __graft_entry__.build()')")
    # torch ships its own copy of the HIP runtime under a different file name
    # (same SONAME).  Loading torch first lets the dynamic linker bind this
    # library to that copy; loading this library first would map a second HIP
    # runtime, and torch's streams and events would be foreign objects to it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    runtimes = hip_runtimes_loaded()
    if len(runtimes) > 1:
        raise ImportError(f"several HIP runtimes are mapped in this process: {runtimes}")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(status):
    """Raise AGDError (NonConvergence for AGD_NOCONV) on a non-zero status."""
    if status != AGD_OK:
        msg = lib.agd_last_error().decode(errors="replace")
        if status == AGD_NOCONV:
            raise NonConvergence(status, msg)
        raise AGDError(status, msg)
    return status


def i64(seq):
    arr = (c_int64 * 3)(*([int(s) for s in seq] + [1] * (3 - len(seq))))
    return arr


def i32(seq):
    arr = (c_int32 * 3)(*([int(s) for s in seq] + [1] * (3 - len(seq))))
    return arr


class Solver:
    """A solver handle (agd_solver): the worklist workspace of the solves run
    through it, bound to the current device and to `stream` (a hipStream_t
    as an int, 0 = the null stream).  Handles share no state, so solves
    through different handles may run concurrently on different streams."""

    def __init__(self, stream=0, desc=None):
        h = c_void_p()
        check(lib.agd_solver_create(desc, stream, ctypes.byref(h)))
        self._h = h
        self.stream = stream

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("the solver handle was closed")
        return self._h

    def solve(self, desc, io, stats):
        return lib.agd_solver_solve(self.handle, desc, io, stats)

    def solve_batched(self, desc, ninst, instances, io, stats):
        return lib.agd_solver_solve_batched(self.handle, desc, ninst, instances, io, stats)

    def check_residual(self, desc, io, atol, rtol):
        n = c_int64(0)
        check(lib.agd_solver_check(self.handle, desc, io, atol, rtol, ctypes.byref(n)))
        return int(n.value)

    def check_stale(self, desc, io, atol, rtol, stale_bits):
        """Cells left above their scheme update, marked in stale_bits."""
        n = c_int64(0)
        check(lib.agd_solver_check_stale(self.handle, desc, io, atol, rtol, stale_bits, ctypes.byref(n)))
        return int(n.value)

    def repair(self, desc, io, atol, rtol, max_rounds):
        """Raise the cells of io.values below their scheme update to +inf,
        round after round; returns (violations left, rounds run)."""
        left, rounds = c_int64(0), c_int32(0)
        check(lib.agd_solver_repair(self.handle, desc, io, atol, rtol, max_rounds,
                                    ctypes.byref(left), ctypes.byref(rounds)))
        return int(left.value), int(rounds.value)

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib.agd_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
