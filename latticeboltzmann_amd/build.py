"""Build recipe for ``liblbm_d2q9.so`` (hand-written HIP kernels + C ABI).

Each ``csrc/*.hip`` translation unit is compiled for gfx950 with its own
flags and the objects are linked into one shared library next to this file.

* ``d2q9_kernels.hip`` and ``ghost.hip`` (the bit-exact FP64/FP32 paths) are
  compiled with ``-ffp-contract=off``: numpy never fuses a multiply into an add, so neither
  may the kernels that must reproduce it.
* every unit is compiled with the ``packed-fp32-ops`` target feature switched
  off, so the compiler never emits packed FP32 arithmetic.
* the reduced-precision units are compiled with ``-fno-slp-vectorize``: with
  the SLP vectorizer on, the ROCm 7.2 gfx950 backend aborts on the FP16 step
  kernels ("Cannot select: v5i32 = bitcast", from packing the nine 16-bit
  outputs of a node); the kernels are bandwidth-bound, so nothing is lost.

Usage: ``python -m latticeboltzmann_amd.build [--changed] [-v]`` (or ``build()``
from Python).
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "liblbm_d2q9.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
BUILD_DIR = os.path.join(HERE, "_build")
ARCH = "gfx950"

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# per-unit flags
UNITS = {
    "d2q9_kernels.hip": ["-ffp-contract=off"],
    "d2q9_lowprec.hip": ["-ffp-contract=off", "-fno-slp-vectorize"],
    "d2q9_batch.hip": ["-ffp-contract=off", "-fno-slp-vectorize"],
    "surrogate.hip": ["-ffp-contract=fast"],
    "ghost.hip": ["-ffp-contract=off"],
}


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build " + LIB_NAME)
    return exe


def _stale(obj, src):
    """True when ``obj`` is older than its unit, any shared header or the C ABI header."""
    if not os.path.exists(obj):
        return True
    deps = [src, os.path.join(HERE, "..", "include", "lbm_d2q9.h")]
    deps += [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith((".hpp", ".h"))]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build(verbose=False, jobs=None, changed_only=False):
    """Compile every HIP unit for gfx950 and link ``liblbm_d2q9.so``.

    The units compile concurrently (at most ``jobs`` at a time, default one
    per unit up to the CPU count); a failing unit raises after the others
    have finished.  ``changed_only`` recompiles only units whose object is
    older than the unit or a header (a development shortcut; the default
    compiles everything)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    cmds, objs = [], []
    for unit, flags in UNITS.items():
        src = os.path.join(CSRC, unit)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD_DIR, unit.replace(".hip", ".o"))
        if not changed_only or _stale(obj, src):
            cmds.append([cc, *COMMON, *flags, "-c", src, "-o", obj])
        objs.append(obj)
    jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1))
    failed = None
    running = []
    for cmd in cmds + [None] * jobs:
        while running and (cmd is None or len(running) >= jobs):
            c, p = running.pop(0)
            if p.wait() != 0 and failed is None:
                failed = subprocess.CalledProcessError(p.returncode, c)
        if cmd is not None:
            if verbose:
                print(" ".join(cmd), flush=True)
            running.append((cmd, subprocess.Popen(cmd)))
    if failed is not None:
        raise failed
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, changed_only="--changed" in sys.argv)
    print(LIB_PATH)
