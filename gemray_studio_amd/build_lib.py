"""Builds libgemray_hip.so (gfx950) in-tree with hipcc: every source compiles to an object in
parallel (the trace kernels are split over one translation unit per anisotropy x environment
class), then one link.

Flags: -ffp-contract=off keeps every multiply-add unfused unless the source calls fmaf
(the reference's arithmetic contract); packed FP32 is disabled
(-target-feature -packed-fp32-ops) so no v_pk_*_f32 is emitted.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgemray_hip.so")
SOURCES = ["transport.hip", "trace_iso_studio.hip", "trace_iso_map.hip", "trace_uni_studio.hip",
           "trace_uni_map.hip", "trace_biax_studio.hip", "trace_biax_map.hip", "runtime.hip", "serve.hip",
           "metrics.hip"]
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False, out: str = LIB, defines=(), extra_flags=(), jobs: int = 0) -> str:
    """Compiles every source (jobs processes at once; 0 = the host's CPU count, at most 10) and
    links them into out."""
    objdir = os.path.join(OBJDIR, os.path.basename(out) + ".obj")
    os.makedirs(objdir, exist_ok=True)
    common = [HIPCC, *FLAGS, *extra_flags, *[f"-D{d}" for d in defines]]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [*common, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        return obj, res

    n = jobs or min(len(SOURCES), max(1, min(os.cpu_count() or 1, 10)))
    with ThreadPoolExecutor(max_workers=n) as pool:
        results = list(pool.map(compile_one, SOURCES))
    failed = [r for _, r in results if r.returncode != 0]
    for r in failed:
        sys.stderr.write(r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed compiling libgemray_hip.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in results]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libgemray_hip.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
