"""Build libawcpml.so for gfx950 in-tree (acousticwavecpml.jl_amd/lib/).

    python acousticwavecpml.jl_amd/build.py

The library is linked from several objects compiled in parallel:
  awcpml.hip          plan/batch management, the C ABI, source/receiver and
                      adjoint kernels
  launch_kernels.hip  the stencil kernels, once per (precision, dimension,
                      plan, batch or two-step batch) via -DAWCPML_REAL/-DAWCPML_NDIM/-DAWCPML_BATCH

-ffp-contract=off keeps every multiply and add separately rounded so the
FP64 kernels reproduce the reference's Float64 arithmetic operation for
operation; packed FP32 instructions are disabled for all HIP code.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib", "libawcpml.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
]
NOISE = "packed-fp32-ops"


def _units():
    units = [("awcpml", os.path.join(CSRC, "awcpml.hip"), [])]
    for real in ("float", "double"):
        for ndim in (2, 3):
            for batch in (0, 1, 2):
                name = "launch_%s_%dd_%s" % (real, ndim, ("plan", "batch", "batch2")[batch])
                units.append((name, os.path.join(CSRC, "launch_kernels.hip"),
                              ["-DAWCPML_REAL=" + real, "-DAWCPML_NDIM=%d" % ndim,
                               "-DAWCPML_BATCH=%d" % batch]))
    return units


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (res.stdout + res.stderr).splitlines() if NOISE not in l and l.strip()]
    return res.returncode, msgs


def build(verbose=True, jobs=None):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    tmpdir = tempfile.mkdtemp(prefix="awcpml_build_")
    try:
        cmds = []
        objs = []
        for name, src, defs in _units():
            obj = os.path.join(tmpdir, name + ".o")
            objs.append(obj)
            cmds.append([HIPCC] + FLAGS + defs + ["-c", "-o", obj, src])
        if verbose:
            print("compiling %d objects (%d jobs): %s" % (len(cmds), jobs, " ".join(cmds[0])),
                  flush=True)
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(_run, cmds))
        failed = False
        for cmd, (rc, msgs) in zip(cmds, results):
            if rc != 0:
                failed = True
                sys.stderr.write(" ".join(cmd) + "\n" + "\n".join(msgs) + "\n")
            elif msgs and verbose:
                print("\n".join(msgs))
        if failed:
            raise RuntimeError("hipcc failed building libawcpml.so")
        tmp = OUT + ".tmp"
        link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
        rc, msgs = _run(link)
        if rc != 0:
            sys.stderr.write(" ".join(link) + "\n" + "\n".join(msgs) + "\n")
            raise RuntimeError("linking libawcpml.so failed")
        os.replace(tmp, OUT)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build())
