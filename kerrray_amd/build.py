"""Build ``libkrhip.so`` (the gfx950 kernels and their C ABI) in-tree with hipcc.

``python -m kerrray_amd.build`` compiles it; ``__graft_entry__.build()``
calls :func:`build`. The library is written next to this file so that it
travels with the source tree.

Flags that matter for results:

* ``-ffp-contract=off``: no fused multiply-add is formed, so every rounding
  happens where the reference's scalar kernel rounds
  (src/kerrray/raytracing/backends/numba_kernel.py runs with
  ``fastmath=False``).
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions are generated.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB_NAME = "libkrhip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ARCH = "gfx950"

SOURCES = ["kr_trace.hip", "kr_batch.hip", "kr_mixed.hip", "kr_adjoint.hip"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP backend)")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", INCLUDE,
        "-I", CSRC,
    ]


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compile every HIP source (in parallel) and link ``libkrhip.so``; always recompiles."""
    import tempfile

    hc = hipcc()
    with tempfile.TemporaryDirectory(prefix="krhip_") as tmpdir:
        procs = []
        objs = []
        for src in SOURCES:
            obj = os.path.join(tmpdir, src + ".o")
            cmd = [hc, *flags(), *(extra or []), "-c", "-o", obj, os.path.join(CSRC, src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
            objs.append(obj)
        failed = False
        for src, proc in procs:
            out, err = proc.communicate()
            if proc.returncode != 0:
                failed = True
                sys.stderr.write(f"--- {src}\n{out}{err}")
            elif verbose and err:
                sys.stderr.write(err)
        if failed:
            raise RuntimeError("hipcc failed")
        tmp = LIB_PATH + ".tmp"
        link = [hc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
        proc = subprocess.run(link, capture_output=True, text=True)
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError("hipcc link failed")
        os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, extra=[a for a in sys.argv[1:] if a != "-v"]))
