"""cagr_apply_queries_noalloc against cagr_apply_queries, through the C ABI.

The no-allocation variant runs the same kernels with the queries' Morton
order and sort buffers in the library's per-GPU static scratch instead of the
handle's grow-only workspace, so its outputs must equal the allocating entry
point's bit for bit.  Two calls in a row, the second smaller (the scratch
keeps the first call's contents), the maximum query count, an empty tree
(zero outputs) and the size limit (rejected before any device work).
"""
import ctypes as ct

import numpy as np
import pytest
import torch

from cagr_amd import WindingNumberTreecode, _lib, synth

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
MAX_QUERIES = 1 << 21

_P, _I, _D, _C = ct.c_void_p, ct.c_int64, ct.c_double, ct.c_int
ARGS = [_P, _C, _C, _P, _C, _P, _C, _I, _P, _D, _C, _D, _P, _P]


def say(*a):
    print("[case]", *a, flush=True)


def lib():
    L = _lib.load()
    for name in ("cagr_apply_queries", "cagr_apply_queries_noalloc"):
        f = getattr(L, name)
        f.restype, f.argtypes = _C, ARGS
    return L


def call(L, name, wt, op, prec, mu, q, width, cont, an, widths=None):
    nq = q.shape[0]
    out = torch.full((nq, _lib.OUT_DIM[op]), float("nan"), dtype=mu.dtype, device=DEV)
    dt = _lib.F32 if mu.dtype == torch.float32 else _lib.F64
    qdt = _lib.F32 if q.dtype == torch.float32 else _lib.F64
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    wp = None if widths is None else ct.c_void_p(widths.data_ptr())
    rc = getattr(L, name)(wt._h, _lib.OP[op], {"fp64": 0, "fp32c": 1}[prec],
                          ct.c_void_p(mu.data_ptr()), dt, ct.c_void_p(q.data_ptr()), qdt, nq, wp,
                          float(width), int(cont), float(an), ct.c_void_p(out.data_ptr()), st)
    _lib.check(rc)
    torch.cuda.synchronize()
    return out


def test_noalloc_queries_equal_the_allocating_entry_point():
    L = lib()
    rng = np.random.default_rng(5)
    p = synth.normalize(synth.ellipsoids(60000, 31)).astype(np.float32)
    n = p.shape[0]
    wt = WindingNumberTreecode(torch.from_numpy(p).to(DEV), leaf_size=32)
    mu32 = torch.from_numpy(synth.random_normals(n, 32).astype(np.float32) * 1e-3).to(DEV)
    mu64 = mu32.double()
    cases = [
        # (queries, op, precision, attrs, width, continuous, an, per-query widths)
        (MAX_QUERIES, "A3", "fp32c", mu32, 0.01, True, 1.0, False),
        (70001, "A", "fp64", mu64, 0.004, True, 0.0, True),
        (5, "A21", "fp32c", mu32, 0.02, False, 1.0, False),
    ]
    for nq, op, prec, mu, width, cont, an, per_query in cases:
        say("noalloc", nq, op, prec)
        q = torch.from_numpy(rng.uniform(-1.2, 1.2, (nq, 3)).astype(np.float32)).to(DEV)
        widths = (torch.from_numpy(rng.uniform(0.001, 0.02, nq)).to(DEV, mu.dtype)
                  if per_query else None)
        want = call(L, "cagr_apply_queries", wt, op, prec, mu, q, width, cont, an, widths)
        got = call(L, "cagr_apply_queries_noalloc", wt, op, prec, mu, q, width, cont, an, widths)
        assert torch.isfinite(want).all()
        assert torch.equal(got, want), (nq, op, prec)
    # the same call again, after the smaller ones, still matches
    say("noalloc repeat")
    q = torch.from_numpy(rng.uniform(-1.2, 1.2, (300000, 3)).astype(np.float32)).to(DEV)
    want = call(L, "cagr_apply_queries", wt, "A3", "fp64", mu64, q, 0.01, True, 1.0)
    for _ in range(2):
        got = call(L, "cagr_apply_queries_noalloc", wt, "A3", "fp64", mu64, q, 0.01, True, 1.0)
        assert torch.equal(got, want)


def test_noalloc_empty_tree_and_size_limit():
    L = lib()
    say("noalloc empty / limit")
    wt = WindingNumberTreecode(torch.zeros((0, 3), device=DEV))
    mu = torch.zeros((0, 3), device=DEV)
    q = torch.rand((1000, 3), device=DEV)
    out = call(L, "cagr_apply_queries_noalloc", wt, "A3", "fp32c", mu, q, 0.01, True, 1.0)
    assert torch.equal(out, torch.zeros_like(out))
    h = WindingNumberTreecode(torch.rand((100, 3), device=DEV))
    mu = torch.rand((100, 3), device=DEV)
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.cagr_apply_queries_noalloc(h._h, _lib.OP["A3"], 1, ct.c_void_p(mu.data_ptr()),
                                      _lib.F32, ct.c_void_p(q.data_ptr()), _lib.F32,
                                      MAX_QUERIES + 1, None, 0.01, 1, 1.0,
                                      ct.c_void_p(q.data_ptr()), st)
    assert rc == 1 and b"2^21" in L.cagr_last_error()
