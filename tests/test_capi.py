"""The C ABI library: loads, exports every function include/spice_hip.h declares, reports
model metadata and validates arguments on the host. No GPU compute here; on a machine without
a GPU the simulation entry points must fail loudly (there is no CPU fallback)."""
import ctypes
import os
import re

import numpy as np
import pytest

from spice_amd import Layout, MODELS, Snn, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "spice_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spice_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_model_info_matches_registry():
    L = _lib.lib()
    for name, (mid, nattrs, sattrs) in MODELS.items():
        na, ns = ctypes.c_int32(), ctypes.c_int32()
        nt, st = (ctypes.c_int32 * 16)(), (ctypes.c_int32 * 16)()
        _lib.check(L.spice_model_info(mid, ctypes.byref(na), ctypes.byref(ns), nt, st))
        assert na.value == len(nattrs) and ns.value == len(sattrs)
        for i, (_, dt) in enumerate(nattrs):
            assert nt[i] == (0 if np.dtype(dt).kind == "f" else 1)
    with pytest.raises(ValueError):
        _lib.check(L.spice_model_info(99, None, None, None, None))


def test_version_string():
    assert b"gfx950" in _lib.lib().spice_version()


def test_invalid_arguments_are_rejected_before_touching_the_gpu():
    with pytest.raises(ValueError):
        Snn("no_such_model", Layout(10, 0.1), 1e-4)
    with pytest.raises(ValueError):
        Snn("brunel", Layout(10, 0.1), -1.0)
    with pytest.raises(ValueError):
        Snn("brunel", Layout(10, 0.1), 1e-4, delay=0)
    with pytest.raises(ValueError):
        Snn.from_ell("brunel", np.zeros(10 * 33, np.int32), 33, 1e-4)  # width % 32 != 0


def _gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_gpu_present(), reason="checks the no-GPU behaviour")
def test_no_silent_cpu_fallback_without_gpu():
    with pytest.raises((RuntimeError, MemoryError)):
        Snn("brunel", Layout(100, 0.1), 1e-4)


def test_unsorted_adjacency_rows_are_rejected():
    """The delivery locates a row's targets inside a tile by binary search, so imported rows
    must be sorted; an unsorted row is an invalid argument (no GPU work is started)."""
    rp = np.array([0, 3, 5, 5], np.int64)
    col = np.array([0, 2, 1, 1, 2], np.int32)  # row 0 unsorted
    with pytest.raises(ValueError, match="sorted"):
        Snn.from_csr("brunel", rp, col, 32, 1e-4)
    ell = np.full((3, 32), np.iinfo(np.int32).max, np.int32)
    ell[1, :2] = [2, 0]  # row 1 unsorted
    with pytest.raises(ValueError, match="sorted"):
        Snn.from_ell("brunel", ell.ravel(), 32, 1e-4)
