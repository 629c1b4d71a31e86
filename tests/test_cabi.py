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
    assert ctypes.sizeof(_lib.EikonalDesc) == 248
    assert _lib.EikonalDesc.factor_radius2.offset == 152
    assert _lib.EikonalDesc.bound_active_blocks.offset == 232
    assert _lib.EikonalDesc.atol.offset == 72
    assert _lib.EikonalDesc.nitermax_o.offset == 88
    assert ctypes.sizeof(_lib.EikonalIO) == 64
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
