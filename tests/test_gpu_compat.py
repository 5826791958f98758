"""The reference-API caller (tests/compat/compat_demo.cpp) linked against the
library reproduces the reference's trees and heatmaps."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, nodes_equal
from oracle import binding as ob
from scenes import GOLDEN_SCENES

DEMO = os.path.join(ROOT, "sequoia_amd", "compat_demo")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.gpu
def test_compat_demo_matches_reference(tmp_path, manifest):
    assert os.path.exists(DEMO), "compat_demo is built by python -m sequoia_amd.build"
    name = "clus3000"
    tris = GOLDEN_SCENES[name]()
    tr = manifest["traces"][name]
    (tmp_path / "tris.bin").write_bytes(tris.tobytes())
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "sequoia_amd") + ":" + os.environ.get("LD_LIBRARY_PATH", ""))
    r = subprocess.run([DEMO, str(tmp_path / "tris.bin"), str(tr["w"]), str(tr["h"]), str(tmp_path / "out")],
                       capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for kind in ("recursive", "horizontal", "grid", "grid_sse", "grid_avx"):
        got = np.fromfile(tmp_path / ("out_%s.nodes" % kind), dtype=ob.NODE_DTYPE)
        m = manifest["builds"]["%s/%s" % (name, kind)]
        if "reference_error" in m:
            continue
        assert sha(got) == m["nodes_sha256"], kind
    for kind in ("morton", "morton2"):
        got = np.fromfile(tmp_path / ("out_%s.nodes" % kind), dtype=ob.NODE_DTYPE)
        links = np.stack([got["count"].astype(np.uint32), got["flags"].astype(np.uint32), got["child"]], 1)
        assert sha(links) == manifest["builds"]["%s/%s" % (name, kind)]["links_sha256"], kind
    assert sha(np.fromfile(tmp_path / "out_basic.rgb", np.uint8)) == tr["basic_rgb_sha256"]
    assert sha(np.fromfile(tmp_path / "out_packet8.rgb", np.uint8)) == tr["packet8_rgb_sha256"]
    assert sha(np.fromfile(tmp_path / "out_bundle64.rgb", np.uint8)) == tr["bundle64_rgb_sha256"]


ALG = os.path.join(ROOT, "sequoia_amd", "compat_algorithms")


@pytest.mark.gpu
@pytest.mark.parametrize("key64", [False, True])
def test_compat_radix_sort_matches_reference(tmp_path, key64):
    """radix_sort<Key, T> through the C++ shim (passes on the GPU) equals the
    reference's sort of the golden 20,000 keys with duplicates (stable: equal
    keys keep their input order); 64-bit keys and 4-bit digits against a
    stable argsort of the low 32 key bits."""
    from conftest import golden
    g = golden("radix_sort32")
    keys = g["keys"]
    if key64:
        rng = np.random.default_rng(3)
        k = (rng.integers(0, 1 << 12, 50000, dtype=np.uint64) << np.uint64(40)) | \
            rng.integers(0, 1 << 8, 50000, dtype=np.uint64)
        cases = [(k, 8, np.argsort(k, kind="stable")),
                 (k, 4, np.argsort(k & np.uint64(0xFFFFFFFF), kind="stable"))]
    else:
        cases = [(keys, 8, g["order"])]
    for k, b, want in cases:
        fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
        fin.write_bytes(k.tobytes())
        r = subprocess.run([ALG, "sort64" if key64 else "sort32", str(fin), str(fout), str(b)],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        raw = fout.read_bytes()
        n = len(k)
        got_keys = np.frombuffer(raw[:n * k.itemsize], k.dtype)
        got_idx = np.frombuffer(raw[n * k.itemsize:], np.uint32)
        assert np.array_equal(got_idx, want), (key64, b)
        assert np.array_equal(got_keys, k[want])
