"""CPU checks of the C-ABI boundary: the library loads and exports exactly the
entry points include/elpaso_amd.h declares (no device calls)."""
import os
import re
import subprocess

import pytest

from elpaso_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "elpaso_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(elp_\w+)\s*\(", src)))


def test_header_matches_binding_list():
    assert _declared() == sorted(_lib.EXPORTED_SYMBOLS)


def test_library_exports_every_symbol():
    path = _lib.library_path()
    if not os.path.exists(path):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (elp_\w+)", out))
    missing = set(_declared()) - exported
    assert not missing, missing
    L = _lib.load_library()
    for name in _declared():
        assert hasattr(L, name)


def test_default_opts_without_gpu():
    # a pure host call through the ABI
    o = _lib.default_opts()
    assert o.rtol == 1e-8 and o.block == 32 and o.precision == _lib.PREC_MIXED
