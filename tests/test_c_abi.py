"""The shared libraries load and export every function the public headers declare
(no compute calls: runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "softx-d-21-00014_amd", "lib")
HEADERS = {"xamg_hip.h": "libxamg_hip.so", "xamg_amg.h": "libxamg.so",
           "xamg_c_api.h": "libxamg.so", "xamg_unet.h": "libxamg_hip.so",
           "xamg_io.h": "libxamg.so"}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n not in ("if", "while", "sizeof", "defined")})


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_symbol_is_exported(header):
    lib = C.CDLL(os.path.join(LIB, HEADERS[header]), mode=C.RTLD_GLOBAL)
    names = declared(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_device_is_an_error_code(X):
    # bad arguments are rejected with a negative code and a message, without aborting
    rc = X.hip().xh_spmv(None, 7, None, None, 0)
    assert rc < 0
    assert b"bad arguments" in X.hip().xh_last_error()
