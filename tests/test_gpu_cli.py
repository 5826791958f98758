"""The command line (sequoia <scene.obj> ...) on the GPU writes the same image
files as the reference's command line (main.cpp:302-414) on the same OBJ
file: render.bmp for builder/tracer/resolution choices and benchmark1..3.bmp
for -benchmark, compared by SHA-256 with the files the reference wrote
(tests/golden/make_golden_io.py). Where the reference crashes (its grid,
horizontal and grid_avx hierarchies have no triangle pointer) the render
must equal the library's own trace of that tree."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import sequoia_amd as sq

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "sequoia_amd", "sequoia")


def say(*a):
    print(*a, flush=True)


def sha_bytes(b):
    return hashlib.sha256(b).hexdigest()


with open(os.path.join(GOLDEN, "io_manifest.json")) as _f:
    IO = json.load(_f)


@pytest.fixture(scope="module")
def obj_file(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    path = d / "ico1280.obj"
    sq.write_obj(path, sq.gen_icosphere(1280, 1, 0.1))
    assert sha_bytes(path.read_bytes()) == IO["ico1280_obj_sha256"]
    return path


def run_cli(args, cwd):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, (args, r.stdout[-400:], r.stderr[-400:])
    return r.stdout


@pytest.mark.parametrize("case", range(len(IO["cli"])))
def test_cli_images_match_reference(tmp_path, obj_file, case):
    c = IO["cli"][case]
    args, files, res = c["args"], c["files"], c["result"]
    say("cli", args)
    out = run_cli([str(obj_file)] + args, tmp_path)
    if "reference_error" in res:
        # the reference cannot trace this builder's tree: compare with the library
        kind = args[args.index("-builder") + 1]
        w, h = int(args[args.index("-res") + 1]), int(args[args.index("-res") + 2])
        K = int(args[args.index("packet") + 1])
        tris = sq.gen_icosphere(1280, 1, 0.1)
        with sq.Context(0) as ctx:
            s = ctx.upload(tris)
            b = ctx.build(s, kind)
            want = ctx.trace_camera(b, sq.reference_camera(tris, w, h), heatmap=sq.HEATMAP_PACKET, width=K,
                                    want_hits=False)["rgb"]
        assert (tmp_path / "render.bmp").read_bytes() == sq.encode_bmp(want)
        return
    for line in res["stdout_geometry"]:
        assert line in out
    for f in files:
        assert sha_bytes((tmp_path / f).read_bytes()) == res[f], f


RELINKED = os.path.join(ROOT, "oracle", "_ref", "sequoia_relinked")


@pytest.mark.skipif(not os.path.exists(RELINKED),
                    reason="the reference's main.cpp relinked against the library is built only where the "
                           "reference sources are (oracle/Makefile relink)")
@pytest.mark.parametrize("case", range(len(IO["cli"])))
def test_reference_cli_relinked_against_library(tmp_path, obj_file, case):
    """The reference's own command line (main.cpp compiled unchanged against
    its own headers) linked against libsequoia_amd.so instead of its builder
    and tracer objects writes the reference's images; its builds and traces
    run on the GPU through the library's Builder::* and trace* symbols."""
    c = IO["cli"][case]
    args, files, res = c["args"], c["files"], c["result"]
    say("relinked reference cli", args)
    r = subprocess.run([RELINKED, str(obj_file)] + args, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, (args, r.stdout[-400:], r.stderr[-400:])
    if "reference_error" in res:
        # the reference's own grid tree has no triangle pointer; the library's has
        kind = args[args.index("-builder") + 1]
        w, h = int(args[args.index("-res") + 1]), int(args[args.index("-res") + 2])
        K = int(args[args.index("packet") + 1])
        tris = sq.gen_icosphere(1280, 1, 0.1)
        with sq.Context(0) as ctx:
            s = ctx.upload(tris)
            b = ctx.build(s, kind)
            want = ctx.trace_camera(b, sq.reference_camera(tris, w, h), heatmap=sq.HEATMAP_PACKET, width=K,
                                    want_hits=False)["rgb"]
        assert (tmp_path / "render.bmp").read_bytes() == sq.encode_bmp(want)
        return
    for line in res["stdout_geometry"]:
        assert line in r.stdout
    for f in files:
        assert sha_bytes((tmp_path / f).read_bytes()) == res[f], f
