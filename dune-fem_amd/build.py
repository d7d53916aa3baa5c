"""Build the MI355X library libdunefem_amd.so in-tree with hipcc (gfx950).

Usage: python dune-fem_amd/build.py [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "dunefem_amd")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libdunefem_amd.so")

SOURCES = ["fem_tables.cpp", "fem_blas.hip", "fem_dg.hip", "fem_cg.hip", "fem_newton.cpp",
           "fem_op.hip", "fem_mg.hip", "fem_csr.hip", "fem_noalloc.hip", "fem_solver.hip", "fem_batch.hip", "fem_cnn.hip", "fem_learned.hip",
           "fem_capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled for all device code
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-I", SRC]


# per-source flags: the DG apply keeps the FP64 MFMA results in ordinary
# VGPRs (the accumulation registers would add to the register total and cost
# a wave per SIMD)
SRC_FLAGS = {"fem_dg.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def compile_one(src):
    obj = os.path.join(OBJ_DIR, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + FLAGS + SRC_FLAGS.get(src, []) + lang + ["-c", os.path.join(SRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    warn = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(jobs=8, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for obj, warn in results:
        if verbose and warn:
            print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + [o for o, _ in results]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (p.stdout, p.stderr))
    os.replace(tmp, LIB)
    return LIB


CPP_TESTS = ["inverseoperatortest"]
CPP_TEST_DIR = os.path.join(ROOT, "tests", "cpp")


def build_cpp_tests():
    """C++ drivers of the host facade (include/dunefem_amd.hh): host code
    only, linked against the library with an rpath relative to the binary."""
    out_dir = os.path.join(CPP_TEST_DIR, "_bin")
    os.makedirs(out_dir, exist_ok=True)
    rel = os.path.relpath(OUT_DIR, out_dir)
    outs = []
    for name in CPP_TESTS:
        exe = os.path.join(out_dir, name)
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-Wextra",
               "-I", os.path.join(ROOT, "include"), os.path.join(CPP_TEST_DIR, name + ".cc"),
               "-L", OUT_DIR, "-ldunefem_amd", "-Wl,-rpath,$ORIGIN/" + rel, "-o", exe]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("C++ test build failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
        outs.append(exe)
    return outs


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
