"""The C-ABI library loads without a GPU and exports every declared symbol."""
import os
import re
import subprocess

from conftest import ROOT

import sequoia_amd as sq
from sequoia_amd import _lib


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//.*", "", src)
        for m in re.finditer(r"\b(sq_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    so = _lib.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = declared_functions() - exported
    assert not missing, missing
    assert set(_lib.declared_symbols()) <= exported


def test_library_loads_and_reports_version():
    assert b"gfx950" in sq.lib().sq_version()


def test_no_device_is_an_error_not_a_crash():
    # Without a GPU, context creation must fail with a status code.
    import ctypes
    if os.path.exists("/dev/kfd"):
        return
    h = ctypes.c_void_p()
    rc = sq.lib().sq_ctx_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value


def test_bad_arguments_rejected():
    import ctypes
    assert sq.lib().sq_ctx_create(0, None) == -1
    assert sq.lib().sq_scene_upload(None, None, 0, None) == -1
    assert b"NULL" in sq.lib().sq_last_error() or b"bad" in sq.lib().sq_last_error()
    assert sq.lib().sq_build(None, None, 0, None, ctypes.byref(ctypes.c_void_p())) == -1
