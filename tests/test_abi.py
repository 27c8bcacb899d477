"""The C-ABI library loads without a GPU and exports every symbol that
include/bitz_hip.h declares, with the binding's table covering all of them."""
import os
import re

import pytest

import bitz_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "bitz_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(bitz_[a-z0-9_]+)\s*\(", text))


def test_every_declared_symbol_is_exported():
    lib = L.load()
    names = _declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert _declared() == set(L.SIGNATURES)


def test_error_path_without_device():
    lib = L.load()
    # a null context is rejected before any device call
    rc = lib.bitz_ctx_synchronize(None)
    assert rc == -4
    assert b"null" in lib.bitz_last_error()
    assert lib.bitz_version().startswith(b"bitz_hip gfx950")


def test_job_struct_layout_matches_the_header(tmp_path):
    """bitz_amd.serve's ctypes bitz_job has the C header's size and offsets"""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    import ctypes as C
    from bitz_amd.serve import _JobStruct
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    names = [f[0] for f in _JobStruct._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"bitz_hip.h\"\nint main(){printf(\"%zu\\n\", "
                   "sizeof(bitz_job));" + "".join(f"printf(\"%zu\\n\", offsetof(bitz_job, {n}));" for n in names)
                   + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", inc, "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got[0] == C.sizeof(_JobStruct)
    assert got[1:] == [getattr(_JobStruct, n).offset for n in names]
