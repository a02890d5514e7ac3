"""Build libgbs_hip.so (the HIP C-ABI library) and the C++ drop-in test
programs in-tree with hipcc / g++ for gfx950.

    python -m gemmbitserial_amd.build          # library + C++ tests
    python -m gemmbitserial_amd.build --lib    # library only

Outputs (git-ignored, shipped with the tree):
    gemmbitserial_amd/libgbs_hip.so
    gemmbitserial_amd/bin/test_dropin      (tests/cpp/test_dropin.cpp)
    gemmbitserial_amd/bin/geometry_probe   (tests/cpp/geometry_probe.cpp)
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgbs_hip.so")
BIN = os.path.join(PKG, "bin")
SOURCES = ["gbs_runtime.hip", "gbs_pack.hip", "gbs_popc.hip", "gbs_mfma.hip",
           "gbs_batch.hip", "gbs_serve.hip", "gbs_bits.hip", "gbs_planes.hip"]
CPP_PROGRAMS = {"test_dropin": "tests/cpp/test_dropin.cpp",
                "geometry_probe": "tests/cpp/geometry_probe.cpp"}

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # packed FP32 VALU ops stay off in every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r.stdout


def build_lib(jobs=8):
    os.makedirs(OBJ, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cpp():
    os.makedirs(BIN, exist_ok=True)
    out = []
    for name, src in CPP_PROGRAMS.items():
        path = os.path.join(ROOT, src)
        exe = os.path.join(BIN, name)
        _run(["g++", "-O2", "-std=c++11", "-Wall", f"-I{INCLUDE}", path, "-o", exe + ".tmp",
              f"-L{PKG}", "-lgbs_hip", f"-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM}/lib"])
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise SystemExit(f"hipcc not found at {HIPCC}")
    print(build_lib(a.j))
    if not a.lib:
        for p in build_cpp():
            print(p)


if __name__ == "__main__":
    main()
