"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (CPU test; nothing is called that touches the device)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b((?:magma|magmablas)_\w+)\s*\(", text):
            names.add(m.group(1))
    # static inline helpers in the headers are not exported symbols
    inline = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        for m in re.finditer(r"static inline [^\n]*?\b(magma_\w+)\s*\(", open(h).read()):
            inline.add(m.group(1))
    return sorted(names - inline)


def test_headers_declare_the_path():
    names = declared_functions()
    for must in ("magma_dgetrf_batched", "magma_dgetrs_batched", "magma_dgesv_batched",
                 "magma_dpotrf_batched", "magma_dpotrs_batched", "magma_dposv_batched",
                 "magma_queue_create_internal", "magma_init"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from magma_amd import lib
    if not os.path.exists(lib.LIB_PATH):
        pytest.fail("libmagma_amd.so not built")
    cdll = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(cdll, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_prototypes_resolve():
    from magma_amd import lib
    L = lib.MagmaLibrary()
    for name in lib.PROTOTYPES:
        assert callable(getattr(L, name))
