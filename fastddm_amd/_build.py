"""Build step of the HIP library: ``python -m fastddm_amd._build``.

Compiles ``csrc/*.hip`` for gfx950 into ``libfddm_hip.so`` beside this file.
Packed FP32 instructions are disabled for all device code
(``-target-feature -packed-fp32-ops``).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["fddm_kernels.hip", "fddm_time_p2.hip", "fddm_spatial_p2.hip", "fddm_spatial_mr.hip", "fddm_capi.hip", "fddm_server.hip",
           "fddm_batch.hip", "fddm_azavg.hip", "fddm_edge.hip",
           "fddm_time_long.hip", "fddm_ingest.hip"]
TARGET = os.path.join(HERE, "libfddm_hip.so")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isfile(cand) or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True, jobs: int = 7) -> str:
    """Compile every HIP source (in parallel) and link the shared library."""
    objs = []
    procs = []
    common = [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        "-I", os.path.join(HERE, "csrc"),
        "-Wno-unused-command-line-argument",
    ]
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *common, "-c", os.path.join(HERE, "csrc", src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(obj)
        if len(procs) >= jobs:
            p, s = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError(f"hipcc failed on {s}")
    for p, s in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {s}")
    tmp = TARGET + ".tmp"
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
