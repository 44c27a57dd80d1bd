"""Builds libsrw_amd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m srw_amd.build   (or srw_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsrw_amd.so")
SOURCES = ["sr.hip", "sr_auto.hip", "fft.hip", "fft_gen.hip", "multie.hip", "extract.hip", "optics.hip", "learn.hip", "moments.hip", "resize.hip", "noalloc.hip", "srwlib_api.cpp", "propag_api.cpp", "multie_api.cpp", "moments_api.cpp", "resize_api.cpp", "auto_api.cpp", "other_api.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def build(verbose=False):
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    for s in srcs:
        src = os.path.join(CSRC, s)
        obj = os.path.join(HERE, "_obj", s + ".o")
        lang = ["-x", "hip"]
        cmd = [HIPCC] + FLAGS + lang + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
