"""The C-ABI library loads on a machine without a GPU and exports every
function include/eqbf.h declares (no compute call is made here)."""
import ctypes
import os
import re

import pytest

from eqtlbma_amd import LIB_PATH, REPO_DIR


def _declared():
    names = set()
    for h in os.listdir(os.path.join(REPO_DIR, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(REPO_DIR, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(eqbf_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_header_symbols():
    assert os.path.exists(LIB_PATH), "build the library first (python -m eqtlbma_amd.build)"
    lib = ctypes.CDLL(LIB_PATH)
    names = _declared()
    assert {"eqbf_create", "eqbf_assoc_uvlr", "eqbf_assoc_uvlr_device"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_config_helpers_without_gpu():
    lib = ctypes.CDLL(LIB_PATH)
    lib.eqbf_n_configs.restype = ctypes.c_int64
    lib.eqbf_n_raw.restype = ctypes.c_int64
    lib.eqbf_n_weighted.restype = ctypes.c_int64
    assert lib.eqbf_n_configs(2, 5) == 31
    assert lib.eqbf_n_configs(1, 44) == 44
    assert lib.eqbf_n_raw(1, 44, 50, 50) == 3 * 50 + 44 * 50
    assert lib.eqbf_n_weighted(2, 3) == 3 + 7 + 2
    import numpy as np
    buf = np.zeros(7, np.uint64)
    assert lib.eqbf_config_masks(2, 3, ctypes.c_void_p(buf.ctypes.data)) == 0
    assert list(buf) == [1, 2, 4, 3, 5, 6, 7]
