"""The C ABI library loads and exports every entry point the public header
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from conftest import PKG, ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "voss", "voss.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(voss_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def libpath():
    import vossgpu
    if not os.path.exists(vossgpu.LIB_PATH):
        subprocess.run([sys.executable, os.path.join(PKG, "build.py"), "lib"], check=True)
    return vossgpu.LIB_PATH


def test_header_declares_entry_points():
    fns = header_functions()
    assert "voss_run_plain" in fns and "voss_mr_batch" in fns
    assert len(fns) >= 15


def test_library_exports_header(libpath):
    lib = ctypes.CDLL(libpath)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_binding_covers_header():
    import vossgpu
    assert sorted(vossgpu.EXPORTS) == header_functions()


def test_no_oracle_in_product():
    """The product sources never reference the oracle."""
    for dirpath, _, files in os.walk(PKG):
        for fn in files:
            if fn.endswith((".hip", ".h", ".cpp", ".py")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, fn)
