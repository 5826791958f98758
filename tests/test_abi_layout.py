"""The ctypes mirror's structures have the C header's sizes and field offsets
(a C program compiled against include/sequoia.h reports them)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

from sequoia_amd import _lib as L

PAIRS = {
    "sq_camera": L.Camera, "sq_build_params": L.BuildParams, "sq_trace_params": L.TraceParams,
    "sq_trace_out": L.TraceOut, "sq_tune_result": L.TuneResult, "sq_server_config": L.ServerConfig,
    "sq_request": L.Request, "sq_request_status": L.RequestStatus, "sq_server_stats": L.ServerStats,
    "sq_timings": L.Timings, "sq_node2_root": L.Node2Root, "sq_nodek_root": L.Node2Root,
}
ARRAYS = {"sq_node": L.NODE_DTYPE, "sq_ray": L.RAY_DTYPE, "sq_node2": L.NODE2_DTYPE, "sq_plucker": L.PLUCKER_DTYPE,
          "sq_node4": L.nodek_dtype(4), "sq_node8": L.nodek_dtype(8)}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_struct_layouts_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sequoia.h"', "int main(void) {"]
    for cname, cls in PAIRS.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    for cname in ARRAYS:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for cname, cls in PAIRS.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out["%s.%s" % (cname, f)]) == getattr(cls, f).offset, (cname, f)
    for cname, dt in ARRAYS.items():
        assert int(out[cname]) == np.dtype(dt).itemsize, cname


REF_SRC = "/root/reference/src"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(REF_SRC),
                    reason="needs g++ and the reference's headers")
def test_wide_node_layout_matches_reference(tmp_path):
    """sq_node4 / sq_node8 have the field offsets of the reference's
    template <size_t K> struct Node (aabb.h:26-32) for K = 4 and 8."""
    src = tmp_path / "nodek.cpp"
    src.write_text("\n".join([
        '#include <cstdio>', '#include <cstddef>', '#include "aabb.h"', "int main() {",
        *['std::printf("%d %zu %zu %zu %zu %zu\\n", {k}, sizeof(Node<{k}>), offsetof(Node<{k}>, min), '
          'offsetof(Node<{k}>, flags), offsetof(Node<{k}>, max), offsetof(Node<{k}>, child));'.format(k=k)
          for k in (4, 8)],
        "return 0; }"]))
    exe = tmp_path / "nodek"
    subprocess.run(["g++", "-std=c++17", "-march=x86-64-v3", "-w", "-I", REF_SRC, str(src), "-o", str(exe)],
                   check=True)
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        k, size, omin, oflags, omax, ochild = map(int, line.split())
        dt = L.nodek_dtype(k)
        assert size == dt.itemsize
        assert (omin, oflags, omax, ochild) == (dt.fields["min_x"][1], dt.fields["flags"][1], dt.fields["max_x"][1],
                                                dt.fields["child"][1])


REF_MAIN_OBJ = os.path.join(ROOT, "oracle", "_ref", "main_cli.o")
LIBSQ = os.path.join(ROOT, "sequoia_amd", "libsequoia_amd.so")


def _syms(args):
    r = subprocess.run(["nm", "-C"] + args, capture_output=True, text=True, check=True)
    return r.stdout.splitlines()


@pytest.mark.skipif(not os.path.exists(REF_MAIN_OBJ) or not os.path.exists(LIBSQ),
                    reason="needs the reference's main.cpp compiled by oracle/Makefile (relink) and the library")
def test_reference_main_links_against_library(tmp_path):
    """The reference's main.cpp, compiled against its own headers, links
    against libsequoia_amd.so with no builder.o, tracer.o or algorithms.o:
    every Builder::* and trace* symbol it leaves undefined is defined by the
    library under the same mangled name (vec::vec3f1 passed as the reference
    passes it)."""
    need = [l.split(None, 1)[1] for l in _syms(["-u", REF_MAIN_OBJ])
            if "Builder::" in l or "trace" in l]
    assert len(need) == 13, need  # six builders, trace, trace_packet<1,2,4,8>, trace_bundle<16,64>
    assert all("vec::vec3f1" in s for s in need if "trace" in s)
    have = {l.split(None, 2)[2] for l in _syms(["-D", "--defined-only", LIBSQ]) if len(l.split(None, 2)) == 3}
    missing = [s for s in need if s not in have]
    assert not missing, missing
    exe = tmp_path / "relinked"
    r = subprocess.run(["g++", "-fopenmp", REF_MAIN_OBJ, "-o", str(exe), "-L" + os.path.dirname(LIBSQ),
                        "-lsequoia_amd"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


ALG = os.path.join(ROOT, "sequoia_amd", "compat_algorithms")


@pytest.mark.skipif(not os.path.exists(ALG), reason="compat_algorithms is built by sequoia_amd/build.py")
def test_compat_interleave_known_answers(manifest):
    """bit_interleave_32/64 through the C++ shim against the reference's
    outputs and against the bit-placement rule on random inputs."""
    def run(x, y, z):
        out = subprocess.run([ALG, "interleave", str(x), str(y), str(z)], capture_output=True, text=True,
                             check=True).stdout.split()
        return int(out[0]), int(out[1])
    for (x, y, z, w32), (x6, y6, z6, w64) in zip(manifest["interleave32"], manifest["interleave64"]):
        assert run(x, y, z)[0] == w32
        assert run(x6, y6, z6)[1] == w64
    rng = np.random.default_rng(9)
    for _ in range(8):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 21, 3))
        want64 = sum((((x >> i) & 1) << (3 * i)) | (((y >> i) & 1) << (3 * i + 1)) | (((z >> i) & 1) << (3 * i + 2))
                     for i in range(21))
        assert run(x, y, z)[1] == want64
        xs, ys, zs = x & 1023, y & 1023, z & 1023
        assert run(xs, ys, zs)[0] == sum((((xs >> i) & 1) << (3 * i)) | (((ys >> i) & 1) << (3 * i + 1)) |
                                         (((zs >> i) & 1) << (3 * i + 2)) for i in range(10))
