"""The C-ABI library loads and exports every symbol include/cagr_amd.h
declares; the Python layer refuses to run without a GPU (no CPU fallback).
No compute call is made here."""
import ctypes as ct
import os
import re

import pytest

from cagr_amd import _lib

HEADER = _lib.HEADER


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(cagr_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_expected_entry_points():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_loads_and_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcagr_amd.so not built (run __graft_entry__.build())")
    L = ct.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name
    L = _lib.load()
    assert L.cagr_version() >= 1


def test_argument_errors_are_reported_without_a_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcagr_amd.so not built")
    L = _lib.load()
    h = ct.c_void_p()
    # invalid leaf size / dtype / theta are rejected before any device work
    assert L.cagr_tree_build(None, 0, 0, 15, 0, 0.5, None, ct.byref(h)) == 1
    assert b"leaf_size" in L.cagr_last_error()
    assert L.cagr_tree_build(None, 0, 7, 15, 1, 0.5, None, ct.byref(h)) == 1
    assert b"dtype" in L.cagr_last_error()
    assert L.cagr_tree_build(None, 0, 0, 16, 1, 0.5, None, ct.byref(h)) == 1
    assert b"max_depth" in L.cagr_last_error()
    assert L.cagr_tree_build(None, 0, 0, 15, 1, 0.0, None, ct.byref(h)) == 1
    assert b"theta" in L.cagr_last_error()
    assert L.cagr_tree_build(None, 10, 0, 15, 1, 0.5, None, ct.byref(h)) == 1
    assert b"null points" in L.cagr_last_error()
    assert L.cagr_apply(None, 0, 0, None, 0, None, 0, 1.0, None, None) == 1
    # PGR octree: depth range, min depth, sample count, null points
    fake = ct.c_void_p(16)
    assert L.cagr_pgr_build(fake, 10, 2, 1, None, ct.byref(h)) == 1
    assert b"max_depth" in L.cagr_last_error()
    assert L.cagr_pgr_build(fake, 10, 14, 1, None, ct.byref(h)) == 1
    assert b"max_depth" in L.cagr_last_error()
    assert L.cagr_pgr_build(fake, 10, 8, 9, None, ct.byref(h)) == 1
    assert b"min_depth" in L.cagr_last_error()
    assert L.cagr_pgr_build(fake, 0, 8, 1, None, ct.byref(h)) == 1
    assert b"number of samples" in L.cagr_last_error()
    assert L.cagr_pgr_build(None, 10, 8, 1, None, ct.byref(h)) == 1
    assert b"null points" in L.cagr_last_error()


def test_no_cpu_fallback():
    import torch
    from cagr_amd import WindingNumberTreecode
    from cagr_amd.pgr import PGROctree
    with pytest.raises(RuntimeError):
        WindingNumberTreecode(torch.zeros(10, 3))
    with pytest.raises(ValueError):
        PGROctree(torch.zeros(10, 3))


def test_gauss_recon_op_is_checked_first():
    """An operator other than the two GaussRecon drivers' fields (A3 of the
    CUDA driver, A_cpu of the CPU driver) is refused before any device work."""
    import torch
    from cagr_amd.pgr import gauss_recon, gauss_recon_values
    p = torch.zeros(10, 3)
    for bad in ("G", "A", "A2"):
        with pytest.raises(ValueError, match="op must be"):
            gauss_recon_values(p, p, op=bad)
        with pytest.raises(ValueError, match="op must be"):
            gauss_recon(p, p, None, op=bad)
