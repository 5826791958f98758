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
