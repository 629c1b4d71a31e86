"""Build recipe of the HIP library libagd_hip.so (gfx950).

Each HIP translation unit under csrc/ is compiled with hipcc for gfx950 and
the objects are linked into one shared library placed next to this file.
Packed-FP32 instruction selection is disabled (-packed-fp32-ops).
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_NAME = "libagd_hip.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
SOURCES = ["layout.hip", "eikonal.hip", "adjoint.hip", "eik_isotropic.hip", "eik_diagonal.hip",
           "eik_riemann.hip", "eik_rander.hip", "eik_tti.hip", "eik_curvature.hip", "surrogate.hip",
           "geodesic.hip", "chart.hip"]
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}")
    return res.stdout


def build(verbose=False, jobs=None):
    """Compile every HIP source and link libagd_hip.so (always rebuilds)."""
    build_dir = os.path.join(PKG_DIR, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs = [os.path.join(build_dir, s.replace(".hip", ".o")) for s in SOURCES]

    def compile_one(pair):
        src, obj = pair
        return _run([HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])

    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        logs = list(ex.map(compile_one, zip(SOURCES, objs)))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_PATH)
    if verbose:
        for log in logs:
            if log.strip():
                print(log)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
