"""Builds libsequoia_amd.so (HIP, gfx950) in-tree, plus the CPU oracle.

Usage: python -m sequoia_amd.build [--jobs N]

Every HIP source is compiled with exact single-precision semantics:
no contraction (fused multiply-adds are explicit), correctly rounded
division and square root, denormals preserved, and packed-FP32 code
generation disabled.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsequoia_amd.so")
ARCH = os.environ.get("SEQUOIA_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def build_library(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hip, cpp = sources()
    cmds = []
    objs = []
    for f in hip:
        o = os.path.join(OBJ, f + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    for f in cpp:
        o = os.path.join(OBJ, f + ".o")
        objs.append(o)
        cmds.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    # the reference-API demo caller used by the drop-in test
    demo = os.path.join(HERE, "compat_demo")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "compat", "compat_demo.cpp"), "-o", demo + ".tmp",
          "-L" + HERE, "-lsequoia_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(demo + ".tmp", demo)
    alg = os.path.join(HERE, "compat_algorithms")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "compat", "compat_algorithms.cpp"), "-o", alg + ".tmp",
          "-L" + HERE, "-lsequoia_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(alg + ".tmp", alg)
    # the reference's command-line harness (sequoia <scene.obj> [-builder ...] ...)
    cli = os.path.join(HERE, "sequoia")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
          os.path.join(HERE, "cli", "sequoia.cpp"), "-o", cli + ".tmp",
          "-L" + HERE, "-lsequoia_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(cli + ".tmp", cli)
    return LIB


def build_oracle(with_reference=True):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if with_reference and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
        if os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "relink"], check=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build_library(a.jobs, a.v))
    if not a.no_oracle:
        build_oracle()


if __name__ == "__main__":
    sys.exit(main())
