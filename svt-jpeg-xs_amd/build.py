"""Builds libsvtjpegxs_amd.so (gfx950) in place with hipcc.

    python svt-jpeg-xs_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsvtjpegxs_amd.so")
ARCH = os.environ.get("JXS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["jxs_kernels.hip", "jxs_enc_ordered.hip", "jxs_host.cpp", "jxs_codestream.cpp", "jxs_encoder.cpp", "jxs_decoder.cpp", "jxs_dec_kernels.hip",
           "jxs_buffers.cpp", "jxs_server.cpp"]

# Packed-FP32 instructions are disabled for all device code built here.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [_hipcc(), "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
               "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-c", path, "-o", obj] + DEVICE_FLAGS
        if src.endswith(".cpp"):
            cmd += ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("compile failed: " + src)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
