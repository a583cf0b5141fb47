"""Build the HIP library ``libfire_amd.so`` (gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libfire_amd.so"
SOURCES = ["k_graph.hip", "k_embed.hip", "k_gemm.hip", "k_orbdet.hip", "k_jastrow.hip", "k_mcmc.hip", "k_stats.hip", "k_ecp.hip", "k_grad.hip", "fire_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False, jobs: int = 6) -> Path:
    objs = []
    procs = []
    for src in SOURCES:
        obj = CSRC / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [HIPCC, "-c", *FLAGS, "-o", str(obj), str(CSRC / src)]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len([p for p in procs if p[1].poll() is None]) >= jobs:
            procs[0][1].wait()
    for src, p in procs:
        out, _ = p.communicate()
        msgs = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and "warning: feature flag" not in l)
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{msgs}")
        if verbose and msgs.strip():
            print(msgs)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        o.unlink(missing_ok=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
