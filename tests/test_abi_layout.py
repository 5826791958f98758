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
    "sq_timings": L.Timings, "sq_node2_root": L.Node2Root,
}
ARRAYS = {"sq_node": L.NODE_DTYPE, "sq_ray": L.RAY_DTYPE, "sq_node2": L.NODE2_DTYPE, "sq_plucker": L.PLUCKER_DTYPE}


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
