"""The C-ABI library loads and exports every entry point its header declares
(no GPU needed)."""
import ctypes
import os
import re

import pytest

import hyperfox_amd as hf
from conftest import ROOT


def declared(header):
    txt = open(os.path.join(ROOT, 'include', header)).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(hfx_[a-z0-9_]+)\s*\(', txt)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(hf.LIB_PATH)
    names = declared('hfx_hip.h')
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_are_reported_not_raised_through_the_abi():
    import numpy as np
    out = np.zeros(7, dtype=np.int32)
    st = hf.lib.hfx_refel_sizes(4, 1, out.ctypes.data_as(ctypes.c_void_p))
    assert st == -1
    assert b'dimension' in hf.lib.hfx_last_error()
    assert hf.lib.hfx_cubature(3, 11, None, None) == -1


def test_unsupported_gpu_orders_rejected_before_any_device_call():
    v = ctypes.c_longlong()
    assert hf.lib.hfx_element_lds_bytes(3, 6, ctypes.byref(v)) == -1
    assert hf.lib.hfx_element_lds_bytes(2, 11, ctypes.byref(v)) == -1
    assert hf.lib.hfx_element_workspace_bytes(3, 6, ctypes.byref(v)) == -1
    assert hf.lib.hfx_element_lds_bytes(3, 3, ctypes.byref(v)) == 0 and v.value > 0
    assert hf.lib.hfx_element_workspace_bytes(3, 3, ctypes.byref(v)) == 0 and v.value == 0


@pytest.mark.parametrize('dim,p', [(3, 4), (3, 5), (2, 7), (2, 8), (2, 9), (2, 10)])
def test_high_orders_keep_the_element_in_a_device_workspace(dim, p):
    # 3-D p = 4, 5 and 2-D p >= 7 (ReferenceElement.cpp:620-623): the
    # element's working set exceeds the 160 KB of LDS, so the kernel takes a
    # per-workgroup slab of device memory instead (2-D p = 6 still fits)
    v = ctypes.c_longlong()
    assert hf.lib.hfx_element_lds_bytes(dim, p, ctypes.byref(v)) == 0 and v.value == 0
    assert hf.lib.hfx_element_workspace_bytes(dim, p, ctypes.byref(v)) == 0 and v.value > 160 * 1024
    assert hf.lib.hfx_element_lds_bytes(2, 6, ctypes.byref(v)) == 0 and 0 < v.value <= 160 * 1024
