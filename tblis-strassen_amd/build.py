"""Build libtblis_strassen_amd.so (HIP kernels + C-ABI) for gfx950 with plain hipcc lines.

Usage: python tblis-strassen_amd/build.py [--jobs N]
The library is written next to this file; object files go to tblis-strassen_amd/_build/.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtblis_strassen_amd.so")

SOURCES = [
    "plan.cpp",
    "driver.cpp",
    "api.cpp",
    "debug.cpp",
    "autotune.cpp",
    "batch.cpp",
    "noalloc.cpp",
    "batch_persist.hip",
    "leaf_gemm.cpp",
    "leaf_f64_multi.hip",
    "leaf_f64_store.hip",
    "leaf_f64_atomic.hip",
    "leaf_f32_multi.hip",
    "leaf_f32_store.hip",
    "aux_kernels.hip",
    "lowp_kernels.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MISTRA_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU ops are kept off for every device pass (hipcc warns on the host pass)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def up_to_date(src, obj):
    """The object is newer than its source and every header of csrc/ and include/ (a coarse rule)."""
    if not os.path.exists(obj):
        return False
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(d, f) for d in (CSRC, os.path.join(ROOT, "include"))
                                         for f in os.listdir(d) if f.endswith((".hpp", ".h"))]
    return all(os.path.getmtime(x) < t for x in deps)


INCREMENTAL = False


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if INCREMENTAL and up_to_date(src, obj):
        return src, 0, "", obj
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    return src, r.returncode, out, obj


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, out, obj in ex.map(compile_one, SOURCES):
            if out.strip() and (verbose or rc != 0):
                print(f"[{src}]\n{out}", file=sys.stderr)
            if rc != 0:
                raise RuntimeError(f"compilation of {src} failed")
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true", help="skip objects newer than their sources and headers")
    a = ap.parse_args()
    INCREMENTAL = a.incremental
    print(build(a.jobs, a.verbose))
