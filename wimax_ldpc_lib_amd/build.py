"""Build recipe for libwimax_ldpc_amd.so (gfx950 HIP kernels + C ABI + C++ API).

    python -m wimax_ldpc_lib_amd.build          # incremental
    python -m wimax_ldpc_lib_amd.build --force  # rebuild everything

Objects go to wimax_ldpc_lib_amd/_build/, the library to
wimax_ldpc_lib_amd/lib/libwimax_ldpc_amd.so, the drop-in command-line tools
(test_encoder, test_decoder) to wimax_ldpc_lib_amd/bin/ and the C++ check
program of the tests to tests/cpp/_bin/ (all in-tree, git-ignored).
"""
import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libwimax_ldpc_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC", "-Wall",
    # IEEE semantics for the bit-exact FP64 message step
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    # no packed FP32 instructions anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include")]

# longest compiles first, so that a parallel build keeps every worker busy
HIP_SOURCES = ["decode_r0_lo.hip", "decode_r0_hi.hip", "decode_r1_lo.hip", "decode_r1_hi.hip",
               "decode_r2_lo.hip", "decode_r2_hi.hip", "decode_r3_lo.hip", "decode_r3_hi.hip",
               "decode_r4_lo.hip", "decode_r4_hi.hip", "decode_r5_lo.hip", "decode_r5_hi.hip",
               "decode_g0.hip", "decode_g1.hip", "decode_g2.hip", "decode_g3.hip", "decode_g4.hip",
               "capi.hip", "mixed.hip", "learned.hip", "trainer.hip", "encoder.hip", "generic.hip", "qc_host.hip",
               "decode_p0.hip", "decode_p1.hip", "decode_p2.hip", "decode_p3.hip", "decode_p4.hip", "decode_p5.hip"]
CXX_SOURCES = ["ldpc_classes.cpp", "matrices.cpp"]
BINDIR = os.path.join(PKG, "bin")
# host programs linked against the library: (source, output, rpath relative to the output)
PROGRAMS = [
    (os.path.join(PKG, "cli", "test_encoder.cpp"), os.path.join(BINDIR, "test_encoder"), "$ORIGIN/../lib"),
    (os.path.join(PKG, "cli", "test_decoder.cpp"), os.path.join(BINDIR, "test_decoder"), "$ORIGIN/../lib"),
    (os.path.join(PKG, "cli", "BER_example.cpp"), os.path.join(BINDIR, "BER_example"), "$ORIGIN/../lib"),
    (os.path.join(PKG, "cli", "Capacity_example.cpp"), os.path.join(BINDIR, "Capacity_example"), "$ORIGIN/../lib"),
    (os.path.join(ROOT, "tests", "cpp", "drop_in_check.cpp"), os.path.join(ROOT, "tests", "cpp", "_bin", "drop_in_check"),
     "$ORIGIN/../../../wimax_ldpc_lib_amd/lib"),
    (os.path.join(ROOT, "tests", "cpp", "matrices_check.cpp"), os.path.join(ROOT, "tests", "cpp", "_bin", "matrices_check"),
     "$ORIGIN/../../../wimax_ldpc_lib_amd/lib"),
]


_INC = re.compile(r'^\s*#\s*include\s*"([^"]+)"', re.M)


def _deps(src, seen=None):
    """Files a source includes with #include "...", transitively (resolved
    against its own directory and include/)."""
    seen = set() if seen is None else seen
    try:
        text = open(src).read()
    except OSError:
        return seen
    for name in _INC.findall(text):
        for d in (os.path.dirname(src), os.path.join(ROOT, "include")):
            p = os.path.normpath(os.path.join(d, name))
            if os.path.exists(p):
                if p not in seen:
                    seen.add(p)
                    _deps(p, seen)
                break
    return seen


def _stale(obj, src, extra=()):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src] + sorted(_deps(src)) + list(extra))


def _compile(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed:\n%s\n%s%s" % (" ".join(cmd), p.stdout, p.stderr))
    return p.stderr


def build(force=False, jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s + ".o")
        objs.append(obj)
        if force or _stale(obj, src):
            cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    for s in CXX_SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, s + ".o")
        objs.append(obj)
        if force or _stale(obj, src):
            cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_compile, cmds):
            if verbose and out.strip():
                sys.stderr.write(out)
    # relink also when an object is newer than the library (a build cut off
    # between its compiles and the link leaves the objects current)
    if force or cmds or not os.path.exists(LIB) or any(os.path.getmtime(o) > os.path.getmtime(LIB) for o in objs):
        tmp = LIB + ".tmp"
        _compile([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
        os.replace(tmp, LIB)
    for src, out, rpath in PROGRAMS:
        if not os.path.exists(src):
            continue
        if force or _stale(out, src, [LIB]):
            os.makedirs(os.path.dirname(out), exist_ok=True)
            _compile(["g++"] + CXX_FLAGS + ["-fopenmp", src, "-o", out, "-L" + LIBDIR, "-lwimax_ldpc_amd",
                                            "-Wl,-rpath," + rpath, "-Wl,-rpath-link,/opt/rocm/lib"])
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.j, verbose=a.v))


if __name__ == "__main__":
    main()
