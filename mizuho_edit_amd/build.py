"""Build libwpf_hip.so in-tree for gfx950 with hipcc (no JIT, no cache).

The library is one shared object made of the HIP translation units under
csrc/.  Every unit is compiled with multiply-add contraction off (the FP64
kernels keep the reference's rounding steps; the FP32 kernels opt back in
locally) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwpf_hip.so")
SOURCES = ["wpf_api.hip", "k_fields.hip", "k_solver.hip", "k_ensemble.hip", "k_learned.hip", "wpf_io.cpp"]
ARCH = os.environ.get("WPF_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    # keep warnings visible but drop the host pass's note about the feature flag
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if msg:
        sys.stderr.write(msg + "\n")


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs = []
    procs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(HERE, "_build", os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmd = ["hipcc", *FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        msg = "\n".join(l for l in (out + err).splitlines()
                        if "packed-fp32-ops" not in l and l.strip())
        if msg:
            sys.stderr.write(msg + "\n")
        if p.returncode != 0:
            failed = True
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
