"""Scene files on the CPU: the OBJ loader reads what the reference's
loadOBJFromFile reads (main.cpp:108-179), the BMP encoder writes the bytes of
the reference's stb writer (main.cpp:335-336), both against fixtures the
reference wrote (tests/golden/make_golden_io.py); plus the defined results
for inputs the reference cannot read, and the command line's argument errors
(parsed before any GPU call)."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import sequoia_amd as sq

CLI = os.path.join(ROOT, "sequoia_amd", "sequoia")


def sha_bytes(b):
    return hashlib.sha256(b).hexdigest()


@pytest.fixture(scope="module")
def io_manifest():
    with open(os.path.join(GOLDEN, "io_manifest.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def io_arrays():
    return np.load(os.path.join(GOLDEN, "ref_io.npz"))


def test_obj_grammar_matches_reference(io_manifest, io_arrays):
    path = os.path.join(GOLDEN, "obj", "quirks.obj")
    with open(path, "rb") as f:
        assert sha_bytes(f.read()) == io_manifest["quirks_obj_sha256"]
    tris, nverts = sq.load_obj(path)
    want = io_arrays["quirks_tris"]
    assert tris.shape == want.shape
    assert np.array_equal(tris.view(np.uint32), want.view(np.uint32))
    assert nverts == 8


def test_written_obj_reads_back_exactly(tmp_path, io_manifest):
    tris = sq.gen_icosphere(1280, 1, 0.1)
    path = tmp_path / "ico1280.obj"
    sq.write_obj(path, tris)
    assert sha_bytes(path.read_bytes()) == io_manifest["ico1280_obj_sha256"]
    back, nverts = sq.load_obj(path)
    assert nverts == 3 * len(tris)
    assert np.array_equal(back.view(np.uint32), tris.view(np.uint32))


@pytest.mark.parametrize("name", ["bmp_23x37", "bmp_96x160"])
def test_bmp_bytes_match_reference(tmp_path, io_manifest, io_arrays, name):
    rgb = io_arrays[name + "_rgb"]
    data = sq.encode_bmp(rgb)
    assert sha_bytes(data) == io_manifest[name + "_sha256"]
    h, w = rgb.shape[:2]
    pad = (-3 * w) % 4
    assert len(data) == 54 + (3 * w + pad) * h
    # first stored row is memory row 0, pixels as B, G, R
    assert data[54:57] == bytes(rgb[0, 0, ::-1])
    sq.write_bmp(tmp_path / "x.bmp", rgb)
    assert (tmp_path / "x.bmp").read_bytes() == data


def write(tmp_path, text, name="t.obj"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return p


def test_obj_defined_results_for_unreadable_inputs(tmp_path):
    # last record without a newline: read (the reference dereferences a null token)
    t, nv = sq.load_obj(write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3"))
    assert t.shape == (1, 9) and nv == 3
    # an empty fourth index (trailing blank) is no index
    t, _ = sq.load_obj(write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3 \n"))
    assert t.shape == (1, 9)
    # indices that name no vertex, missing fields, missing file: errors
    for text in ["v 0 0 0\nf 1 2 3\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "v 0 0 0\nf -4 1 1\n",
                 "v 1 2\n", "v 0 0 0\nf 1 1\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1  2 3\n"]:
        with pytest.raises(sq.SqError):
            sq.load_obj(write(tmp_path, text))
    with pytest.raises(sq.SqError):
        sq.load_obj(tmp_path / "missing.obj")


def test_obj_empty_tokens_and_late_vertices(tmp_path):
    # two blanks before a coordinate: an empty token reads as 0 and shifts the rest
    t, _ = sq.load_obj(write(tmp_path, "v  5 6 7\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"))
    assert t[0, :3].tolist() == [0.0, 5.0, 6.0]
    # faces may name vertices defined later (triangles are assembled at the end)
    t, _ = sq.load_obj(write(tmp_path, "f 1 2 3\nv 0 0 0\nv 1 0 0\nv 0 1 0\n"))
    assert t.shape == (1, 9) and t[0, 3] == 1.0


def run_cli(args, cwd):
    return subprocess.run([CLI] + args, capture_output=True, text=True, cwd=cwd, timeout=60)


def test_cli_argument_errors(tmp_path):
    assert os.path.exists(CLI), "the command line is built by python -m sequoia_amd.build"
    r = run_cli([], tmp_path)
    assert r.returncode == 0 and "Error: not enough arguments" in r.stdout and "Usage:" in r.stdout
    r = run_cli(["x.obj", "-builder", "quadtree"], tmp_path)
    assert r.returncode == 0 and "Error: invalid builder type" in r.stdout
    r = run_cli(["x.obj", "-tracer", "cone"], tmp_path)
    assert r.returncode == 0 and "Error: invalid tracer type" in r.stdout
    r = run_cli(["x.obj", "-tracer", "packet", "16"], tmp_path)
    assert r.returncode == 1 and "invalid packet width" in r.stderr
    r = run_cli(["x.obj", "-tracer", "bundle", "32"], tmp_path)
    assert r.returncode == 1 and "invalid bundle size" in r.stderr
    r = run_cli(["x.obj", "-res", "64"], tmp_path)
    assert r.returncode == 1 and "needs a value" in r.stderr
