"""Builds the native parts of tb8gen_amd in-tree.

  libtbg.so   -- HIP kernels + C ABI (include/tbg.h), gfx950 code objects, and the
                 slice-file codec (zstd)
  tb8gen      -- the C host driver (tb8gen_main.c) linked against libtbg.so

Run: python -m tb8gen_amd.build   (or __graft_entry__.build()).
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libtbg.so")
DRIVER = os.path.join(PKG, "tb8gen")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TBG_ARCH", "gfx950")

HIP_SOURCES = ["tbg_kernels.hip", "tbg_host.hip"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # never let the compiler form packed fp32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# host-only sources (the slice-file codec), built with g++ against zstd: the
# header from the conda tree, the library the system's (its stable API only;
# linked by path so no other library directory enters the search)
CXX_SOURCES = ["tbg_io.cpp"]
ZSTD_INC = os.environ.get("TBG_ZSTD_INC", "/opt/conda/include")
ZSTD_LIB = os.environ.get("TBG_ZSTD_LIB", "/usr/lib/x86_64-linux-gnu/libzstd.so.1")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", ZSTD_INC]


def build_lib():
    objs = []
    os.makedirs(os.path.join(PKG, "_obj"), exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(PKG, "_obj", src.replace(".hip", ".o"))
        run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(PKG, "_obj", src.replace(".cpp", ".o"))
        run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, ZSTD_LIB])
    os.replace(tmp, LIB)


def build_driver():
    run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", os.path.join(ROOT, "include"),
         "-o", DRIVER, os.path.join(CSRC, "tb8gen_main.c"),
         "-L", PKG, "-ltbg", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + PKG])


def build():
    build_lib()
    build_driver()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "lib" in what:
        build_lib()
    if "all" in what or "driver" in what:
        build_driver()
