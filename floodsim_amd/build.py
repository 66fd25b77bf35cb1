"""Builds the HIP library and the N-API addon in-tree (no JIT caches, no pip installs).

    python -m floodsim_amd.build          # libfloodsim_hip.so (+ floodsim_napi.node when Node headers exist)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FS_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep hipcc from forming v_pk_*_f32 (plain FP32 VALU only)
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build_hip(extra=()):
    out = os.path.join(HERE, "libfloodsim_hip.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, *extra, os.path.join(HERE, "csrc", "floodsim.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_napi():
    """N-API addon (Node's own headers; links libfloodsim_hip.so next to it)."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")) or shutil.which("gcc") is None:
        return None
    out = os.path.join(HERE, "js", "floodsim_napi.node")
    run(["gcc", "-O2", "-fPIC", "-shared", "-I" + inc, "-I" + os.path.join(ROOT, "include"),
         os.path.join(HERE, "csrc", "napi_addon.c"), "-o", out, "-L" + HERE, "-lfloodsim_hip",
         "-Wl,-rpath,$ORIGIN/.."])
    return out


def build_all():
    build_hip()
    if os.path.exists(os.path.join(HERE, "csrc", "napi_addon.c")):
        build_napi()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "libfloodsim_hip.so"))
