"""The C ABI (no GPU needed): libqed.so loads and exports every function declared in
include/qed.h, and the Python binding declares a signature for each of them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "qed.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(qed_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = header_functions()
    assert len(names) >= 50
    for n in ("qed_sector_create", "qed_hv", "qed_lanczos_gs", "qed_band_lanczos", "qed_batch_solve",
              "qed_cdmft_fit_batch", "qed_mlp_train", "qed_mlp_forward_bf16", "qed_ed_ground_state_solve"):
        assert n in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "pyqcm_amd", "libqed.so")
    if not os.path.exists(path):
        pytest.skip("libqed.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    from pyqcm_amd import _lib
    bound = {n for n, _, _ in _lib.SIGNATURES}
    missing = [n for n in header_functions() if n not in bound]
    assert not missing, missing
