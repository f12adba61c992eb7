"""Builds libpointneighbors_hip.so in-tree for gfx950.

Flags: -ffp-contract=off keeps a*b+c unfused (the reference's arithmetic has no FMA);
packed FP32 instructions are disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
# translation units, compiled in parallel (each instantiates a share of the kernels)
SOURCES = [os.path.join(HERE, "csrc", f) for f in
           ("pn_api.hip", "pn_api_trav.hip", "pn_api_sph.hip", "pn_api_vjp.hip", "pn_api_ext.hip",
            "pn_api_noalloc.hip")]
OUT = os.path.join(HERE, "libpointneighbors_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _run(cmd, verbose):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd[-3:])}")
    if verbose:
        sys.stderr.write(res.stderr)


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    with tempfile.TemporaryDirectory(prefix="pn_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in SOURCES]
        with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            list(ex.map(lambda so: _run([hipcc, *HIPCC_FLAGS, "-c", "-o", so[1], so[0]], verbose),
                        zip(SOURCES, objs)))
        tmp = OUT + ".tmp"
        _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
