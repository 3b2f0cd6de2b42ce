"""Build recipe for libpfdd_gpu.so (gfx950) and the pfdd driver, in-tree.

    python -m pfdd_amd.build            # or __graft_entry__.build()
    python -m pfdd_amd.build --incremental   # recompile only objects older than their sources

hipcc cross-compiles for gfx950 without a GPU.  Every HIP compile line carries
-Xclang -target-feature -Xclang -packed-fp32-ops (packed FP32 ops disabled).
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libpfdd_gpu.so")
DRIVER = os.path.join(PKG, "pfdd")
ARCH = os.environ.get("PFDD_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{INC}", f"-I{CSRC}"]
HIP_SOURCES = ["pfdd_api.hip"]
C_SOURCES = ["pfdd_setup.c"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    return out


def _fresh(obj, src):
    """obj exists and is newer than src and every header it may include."""
    if not os.path.exists(obj):
        return False
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith((".hpp", ".h"))] + \
        [os.path.join(INC, h) for h in os.listdir(INC)]
    t = os.path.getmtime(obj)
    return all(os.path.getmtime(d) < t for d in deps)


def build(verbose=False, incremental=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in C_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        if incremental and _fresh(o, os.path.join(CSRC, s)):
            continue
        _run(["gcc", "-O2", "-fPIC", "-std=c99", "-Wall", f"-I{INC}", "-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        if incremental and _fresh(o, os.path.join(CSRC, s)):
            objs.append(o)
            continue
        out = _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        if verbose and out.strip():
            print(out)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    drv_src = os.path.join(PKG, "driver", "pfdd_main.c")
    if os.path.exists(drv_src):
        _run(["gcc", "-O2", "-std=c99", "-Wall", f"-I{INC}", drv_src, "-o", DRIVER, f"-L{PKG}", "-lpfdd_gpu",
              "-Wl,-rpath,$ORIGIN", "-lm"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
