"""Build of the HIP/C++ library liblbfs_amd.so (in-tree, gfx950).

Every source under csrc/ is compiled with hipcc for gfx950 and linked into
lbfs_amd/liblbfs_amd.so next to this file.  Packed FP32 instructions are
disabled for all device code.
"""
import concurrent.futures
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG_DIR, "liblbfs_amd.so")
# host check programs: the C++ mirror of the solver class and the
# reference-signature step entry, each built from host/<name>.cpp
HOST_CHECKS = ("gpu_solver_check", "reference_entry_check")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC, "-I" + os.path.join(REPO, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, LIB)
    # the C++ host mirror's check programs (plain host C++ against the C ABI)
    host = os.path.join(ROOT, "host")
    exes = []
    for name in HOST_CHECKS:
        exe = os.path.join(host, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + host, "-I" + os.path.join(REPO, "include"),
               os.path.join(host, name + ".cpp"), "-L" + PKG_DIR, "-llbfs_amd",
               "-Wl,-rpath,$ORIGIN/../lbfs_amd", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("g++ failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
        exes.append(exe)
    if verbose:
        print("built", LIB, "and", ", ".join(exes))
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
