"""Build libtnkit_hip.so in-tree with hipcc for gfx950.

    python -m tnkit_amd.build            # compile every csrc/*.hip, link tnkit_amd/libtnkit_hip.so

Objects go to tnkit_amd/csrc/_build/. The library is written next to this file so that the
package loads it from its own directory.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtnkit_hip.so")
ARCH = os.environ.get("TNKIT_HIP_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 are generated (they are avoided on this target).
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; libtnkit_hip.so needs ROCm's hipcc")


def _compile(src: str, obj: str) -> None:
    cmd = [hipcc(), *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n" + "\n".join(msgs))


def build(verbose: bool = True) -> str:
    """Compile all HIP sources and link the shared library. Returns its path."""
    bdir = os.path.join(CSRC, "_build")
    os.makedirs(bdir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(bdir, s[:-4] + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(lambda so: _compile(os.path.join(CSRC, so[0]), so[1]), zip(srcs, objs)))
    tmp = OUT + ".tmp"
    r = subprocess.run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    # the C-ABI driver (host program linked against the library by rpath)
    exe = os.path.join(HERE, "cabi_bench")
    r = subprocess.run([hipcc(), "-O2", "-std=c++17", os.path.join(CSRC, "cabi_bench.cpp"),
                        "-o", exe + ".tmp", "-L", HERE, "-ltnkit_hip", "-Wl,-rpath,$ORIGIN"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cabi_bench build failed:\n" + r.stdout + r.stderr)
    os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources, and {exe}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
