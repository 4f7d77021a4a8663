"""Builds bembel_amd/libbembel_amd.so (HIP for gfx950) and the CPU oracle library.

Usage: python -m bembel_amd.build [--jobs N]
"""
import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbembel_amd.so")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH, "-I" + os.path.join(ROOT, "include"),
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
# host setup code decides admissibility and quadrature degrees: no FP contraction there
SOURCES = {
    "bmi_setup.cpp": ["-ffp-contract=off", "-x", "hip"],
    "bmi_nearfield.hip": [],
    "bmi_fmm.hip": [],
    "bmi_solver.hip": [],
    "bmi_capi.hip": [],
    "bmi_server.hip": [],
    "bmi_space_dev.cpp": ["-x", "hip"],
    "bmi_potential.hip": [],
}


def source_digest():
    """sha256 (first 16 hex digits) over the library's sources and its C header, in name order: the
    library built from them reports the same digest in bmi_version()"""
    h = hashlib.sha256()
    names = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp", ".h", ".inc")))
    paths = [os.path.join(CSRC, f) for f in names] + [os.path.join(ROOT, "include", "bembel_amd.h")]
    for path in paths:
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _compile(src, extra, digest):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ['-DBMI_SOURCE_DIGEST="%s"' % digest, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr[-8000:]))
    return obj


def build_product(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    digest = source_digest()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, e, digest) for s, e in SOURCES.items()]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-8000:])
    os.replace(tmp, LIB)
    return LIB


def build_examples():
    """C++ program written against the reference-style facade (include/bembel_amd/Bembel.hpp)"""
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "facade_example")
    cmd = ["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "facade_example.cpp"), "-L" + HERE, "-lbembel_amd",
           "-Wl,-rpath,$ORIGIN/..", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade example build failed:\n" + r.stderr[-4000:])
    return out


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", odir, "-j4"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout[-4000:] + r.stderr[-4000:])
    return os.path.join(odir, "liboracle_bembel.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 4))
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    print("built", build_product(a.jobs))
    print("built", build_examples())
    if not a.no_oracle:
        print("built", build_oracle())


if __name__ == "__main__":
    main()
