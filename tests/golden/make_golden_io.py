"""Writes the scene-file fixtures (OBJ reading, BMP bytes, command-line
renders) from the reference's own code, compiled by oracle/Makefile into
oracle/_ref/libseqref.so (main.cpp's loader, stb BMP writer and main()).

    python tests/golden/make_golden_io.py

Outputs: tests/golden/ref_io.npz and tests/golden/io_manifest.json. The
reference command line ends with exit(0), so each run is a child process in
a scratch directory; its image files are hashed and the directory removed.
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.dirname(os.path.abspath(__file__))
QUIRKS = os.path.join(OUT, "obj", "quirks.obj")

# (args after the OBJ path, image files written). The reference's horizontal,
# grid and grid_avx builders leave Hierarchy::tri_base unset (builder.cpp
# build_hierarchy_horizontal/_grid/_grid_m128), so its command line crashes
# when it traces their trees; one such case is kept to record that. Its
# bundle tracer writes past the image unless both sides are multiples of 8.
CLI_CASES = [
    (["-builder", "recursive", "-tracer", "basic", "-res", "160", "96"], ["render.bmp"]),
    (["-builder", "grid_sse", "-tracer", "packet", "8", "-res", "160", "96"], ["render.bmp"]),
    (["-builder", "recursive", "-tracer", "bundle", "64", "-res", "160", "96"], ["render.bmp"]),
    (["-builder", "recursive", "-tracer", "basic", "-res", "161", "97"], ["render.bmp"]),
    (["-builder", "grid_sse", "-tracer", "packet", "2", "-res", "64", "36"], ["render.bmp"]),
    (["-builder", "recursive", "-tracer", "bundle", "16", "-threads", "4"], ["render.bmp"]),
    (["-builder", "grid", "-tracer", "packet", "8", "-res", "160", "96"], ["render.bmp"]),
    (["-benchmark"], ["benchmark1.bmp", "benchmark2.bmp", "benchmark3.bmp"]),
]


def sha_bytes(b):
    return hashlib.sha256(b).hexdigest()


def run_reference_cli(obj, args, files):
    from oracle import binding as ob
    d = tempfile.mkdtemp(prefix="golden_cli_")
    try:
        code = ("import ctypes, sys\n"
                "l = ctypes.CDLL(%r)\n"
                "a = [s.encode() for s in sys.argv[1:]]\n"
                "argv = (ctypes.c_char_p * (len(a) + 1))(*a, None)\n"
                "l.ref_cli_main(len(a), argv)\n") % ob.REF_SO
        r = subprocess.run([sys.executable, "-c", code, "sequoia", obj] + args, cwd=d, capture_output=True,
                           text=True)
        if r.returncode != 0:
            return {"reference_error": "exit status %d %s" % (r.returncode, (r.stderr or "")[-160:])}
        out = {"stdout_geometry": [l.strip() for l in r.stdout.splitlines() if "faces:" in l or "verts:" in l]}
        for f in files:
            with open(os.path.join(d, f), "rb") as fh:
                out[f] = sha_bytes(fh.read())
        return out
    finally:
        shutil.rmtree(d)


def main():
    from oracle import binding as ob
    import sequoia_amd as sq
    man = {}
    arrays = {}
    # OBJ grammar fixture
    q = ob.ref_load_obj(QUIRKS)
    arrays["quirks_tris"] = q
    with open(QUIRKS, "rb") as fh:
        man["quirks_obj_sha256"] = sha_bytes(fh.read())
    # BMP bytes for an odd width (row padding) and a multiple-of-4 width
    rng = np.random.default_rng(77)
    for name, (h, w) in {"bmp_23x37": (23, 37), "bmp_96x160": (96, 160)}.items():
        rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        path = os.path.join(tempfile.gettempdir(), "golden_%s.bmp" % name)
        ob.ref_write_bmp(path, rgb)
        with open(path, "rb") as fh:
            data = fh.read()
        os.remove(path)
        arrays[name + "_rgb"] = rgb
        man[name + "_sha256"] = sha_bytes(data)
    # command line on a generated icosphere written as OBJ by the library
    tris = sq.gen_icosphere(1280, 1, 0.1)
    d = tempfile.mkdtemp(prefix="golden_obj_")
    try:
        obj = os.path.join(d, "ico1280.obj")
        sq.write_obj(obj, tris)
        with open(obj, "rb") as fh:
            man["ico1280_obj_sha256"] = sha_bytes(fh.read())
        back = ob.ref_load_obj(obj)
        assert np.array_equal(back.view(np.uint32), tris.view(np.uint32)), "OBJ text does not read back exactly"
        man["cli"] = []
        for args, files in CLI_CASES:
            res = run_reference_cli(obj, args, files)
            man["cli"].append({"args": args, "files": files, "result": res})
            print(args, res, flush=True)
    finally:
        shutil.rmtree(d)
    np.savez_compressed(os.path.join(OUT, "ref_io.npz"), **arrays)
    with open(os.path.join(OUT, "io_manifest.json"), "w") as f:
        json.dump(man, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
