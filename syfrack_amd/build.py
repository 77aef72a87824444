"""Build the in-tree HIP library syfrack_amd/libsyfrack_hip.so for gfx950.

hipcc cross-compiles without a GPU.  Flags:
  -ffp-contract=off        every expression rounds as written (parity with the oracle's
                           f32/f64 restatement of the reference arithmetic)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                           no packed FP32 instructions are generated
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsyfrack_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


OBJ_DIR = os.path.join(HERE, "_build")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    err = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if verbose and err:
        sys.stderr.write("\n".join(err) + "\n")


def build(verbose: bool = False, extra=None, jobs: int = 0) -> str:
    """Compiles every source to an object in parallel (each file is its own code object; the
    host launchers sit beside their kernels), then links the shared library."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ_DIR, os.path.basename(s)[:-4] + ".o") for s in srcs]
    compile_flags = [f for f in FLAGS if f != "-shared"] + (extra or [])
    cmds = [[HIPCC] + compile_flags + ["-c", s, "-o", o] for s, o in zip(srcs, objs)]
    workers = jobs or min(len(cmds), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    # hiprtc compiles the kernels of specialised graph topologies at run time (graph.hip)
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-L/opt/rocm/lib", "-lhiprtc", "-o", OUT + ".tmp"],
         verbose)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
