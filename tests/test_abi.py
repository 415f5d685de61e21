"""The C ABIs: the product library loads without a GPU and exports every
function the public C headers declare (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["include/spoa_hip.h", "include/spoa_capi.h"]
LIB = os.path.join(ROOT, "spoa_amd", "libspoa_amd.so")


def declared_functions(path):
    text = open(os.path.join(ROOT, path)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    names = set()
    # a declaration: return type, name, '(' ... ');' at file scope
    for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
        name = m.group(1)
        if name.startswith("spoa_"):
            names.add(name)
    return names


def test_headers_declare_functions():
    for h in HEADERS:
        assert len(declared_functions(h)) >= 10, h


@pytest.mark.skipif(not os.path.exists(LIB), reason="product library not built")
def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(LIB)
    missing = []
    for h in HEADERS:
        for name in sorted(declared_functions(h)):
            try:
                getattr(lib, name)
            except AttributeError:
                missing.append((h, name))
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="product library not built")
def test_last_error_is_callable_without_device():
    lib = ctypes.CDLL(LIB)
    lib.spoa_hip_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.spoa_hip_last_error(), bytes)


@pytest.mark.skipif(not os.path.exists(LIB) or os.path.exists("/dev/kfd"), reason="needs a host without a GPU")
def test_product_fails_loudly_without_device():
    """No CPU fallback: engines and servers refuse to start without a GPU."""
    import sys
    sys.path.insert(0, ROOT)
    import spoa_amd
    with pytest.raises(ValueError):
        spoa_amd.AlignmentEngine.create(spoa_amd.NW, 5, -4, -8)
    with pytest.raises(ValueError):
        spoa_amd.Server(threads=1)
