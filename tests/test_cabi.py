"""The C-ABI library loads on a CPU-only host and exports every function
declared in include/*.h; the ctypes signature table covers them all."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(agd_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"agd_eikonal_solve", "agd_eikonal_flow", "agd_eikonal_adjoint",
            "agd_block_expand", "agd_prepare_geometry", "agd_stencil_export"} <= names


def test_library_exports_every_declared_symbol():
    from adaptivegriddiscretizations_amd import _lib
    lib = ctypes.CDLL(_lib._build.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == declared_functions()


def test_struct_layouts_match_header():
    from adaptivegriddiscretizations_amd import _lib
    # offsets follow the natural alignment of the C declarations
    assert ctypes.sizeof(_lib.EikonalDesc) == 320
    assert _lib.EikonalDesc.factor_radius2.offset == 152
    assert _lib.EikonalDesc.bound_active_blocks.offset == 296
    assert _lib.EikonalDesc.atol.offset == 72
    assert _lib.EikonalDesc.nitermax_o.offset == 88
    assert ctypes.sizeof(_lib.EikonalIO) == 80
    assert _lib.EikonalIO.scheme_codes_len.offset == 72
    assert ctypes.sizeof(_lib.EikonalStats) == 56
    assert ctypes.sizeof(_lib.BatchInstance) == 64


def test_error_paths_without_device():
    """Argument checks run before any device call and report through
    agd_last_error."""
    from adaptivegriddiscretizations_amd import _lib
    d = _lib.EikonalDesc()
    d.ndim = 4
    st = _lib.lib.agd_eikonal_solve(d, None, None, None)
    assert st == _lib.AGD_EINVAL
    assert b"ndim" in _lib.lib.agd_last_error()
    assert _lib.lib.agd_tile_supported(_lib.RIEMANN, 3, _lib.i32((4, 4, 4))) == 1
    assert _lib.lib.agd_tile_supported(_lib.RIEMANN, 3, _lib.i32((3, 3, 3))) == 0


def test_single_hip_runtime_in_fresh_process():
    """Loading the library must not map a second HIP runtime next to torch's
    (torch's stream handles would be foreign objects to it)."""
    import subprocess
    import sys
    code = ("from adaptivegriddiscretizations_amd import _lib; import torch; "
            "print(len(_lib.hip_runtimes_loaded()))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"


def test_scheme_lengths_and_refused_short_buffers():
    """agd_scheme_lengths gives the component counts of the precomputed
    scheme (TTI: 7 sub-schemes, the min/max flag, and with order 2 the
    further-neighbour codes); a solve or scheme export whose buffers are
    shorter than that layout is refused before any device access."""
    from adaptivegriddiscretizations_amd import _lib
    w, c = ctypes.c_int64(), ctypes.c_int64()
    expect = {(_lib.RIEMANN, 2, 1): (3, 6), (_lib.RIEMANN, 3, 1): (6, 12),
              (_lib.TTI, 2, 1): (21, 43), (_lib.TTI, 2, 2): (21, 85),
              (_lib.TTI, 3, 1): (42, 85), (_lib.TTI, 3, 2): (42, 169),
              # curvature models: offsets (forward ones included) and the flag
              (_lib.REEDSSHEPP2, 3, 1): (6, 13), (_lib.REEDSSHEPPFWD2, 3, 1): (7, 15),
              (_lib.ELASTICA2, 3, 1): (30, 61), (_lib.DUBINS2, 3, 1): (12, 25)}
    for (model, d, order), wc_cc in expect.items():
        assert _lib.lib.agd_scheme_lengths(model, d, order, ctypes.byref(w), ctypes.byref(c)) == 0
        assert (w.value, c.value) == wc_cc

    n = 16
    d = _lib.EikonalDesc()
    d.model, d.ndim, d.scalar, d.solver = _lib.TTI, 2, _lib.F32, _lib.AGSI
    for k in range(3):
        d.shape[k] = n if k < 2 else 1
        d.shape_i[k] = 8 if k < 2 else 1
    d.niter_i, d.nitermax_o, d.check_interval, d.order = 8, 10, 1, 2
    size_tot = n * n
    io = _lib.EikonalIO()
    # placeholders: the length check fails before anything is dereferenced
    io.values = io.geom = io.rhs = io.seeds = io.scheme_weights = io.scheme_codes = 256
    io.scheme_weights_len = 21 * size_tot
    io.scheme_codes_len = 43 * size_tot        # sized for order 1
    assert _lib.lib.agd_eikonal_solve(d, io, None, None) == _lib.AGD_EINVAL
    assert b"scheme_codes_len" in _lib.lib.agd_last_error()
    assert _lib.lib.agd_eikonal_scheme(d, io, None) == _lib.AGD_EINVAL
    assert b"scheme_codes_len" in _lib.lib.agd_last_error()
    io.scheme_codes_len = 85 * size_tot
    io.scheme_weights_len = 20 * size_tot      # one component short
    assert _lib.lib.agd_eikonal_solve(d, io, None, None) == _lib.AGD_EINVAL
    assert b"scheme_weights_len" in _lib.lib.agd_last_error()
