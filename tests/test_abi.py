"""The C-ABI library loads without a GPU and exports every symbol that
include/bitz_hip.h declares, with the binding's table covering all of them."""
import os
import re

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
