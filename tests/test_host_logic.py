"""CPU replay of the HIP tree build (tests/host_logic/build_logic.cpp runs the
per-element bodies of cagr_amd/csrc/tree_logic.h serially, in the kernels'
order) against the oracle tree, including degenerate inputs; compiled with
UBSan (-fno-sanitize-recover: any report aborts) and libstdc++'s checked
container indexing (_GLIBCXX_ASSERTIONS), so an out-of-range index in the
shared logic fails here rather than on a GPU."""
import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

import oracle
from cagr_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hl(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("hl") / "libhl.so")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-fsanitize=undefined",
                    "-fno-sanitize-recover=all", "-D_GLIBCXX_ASSERTIONS", "-o", so,
                    os.path.join(HERE, "host_logic", "build_logic.cpp")], check=True)
    L = ct.CDLL(so)
    P, I = ct.c_void_p, ct.c_int64
    L.hl_build.argtypes = [P, I, ct.c_int, ct.c_int, ct.c_double, P, P, P]
    L.hl_build.restype = I
    L.hl_export.argtypes = [P, I, P, I, P] + [P] * 7
    L.hl_export.restype = I
    return L


def hl_tree(L, p, md, S, theta=0.5):
    p = np.ascontiguousarray(p, np.float64)
    n = p.shape[0]
    P = ct.c_void_p
    m = np.zeros(1, np.int64)
    L.hl_build(p.ctypes.data_as(P), n, md, S, theta, m.ctypes.data_as(P), None, None)
    M = int(m[0])
    topo = np.zeros((max(M, 1), 32), np.uint8)
    perm = np.zeros(max(n, 1), np.int32)
    L.hl_build(p.ctypes.data_as(P), n, md, S, theta, m.ctypes.data_as(P),
               topo.ctypes.data_as(P), perm.ctypes.data_as(P))
    t = oracle._alloc_tree(M, 0)
    ll = L.hl_export(p.ctypes.data_as(P), n, topo.ctypes.data_as(P), M, perm.ctypes.data_as(P),
                     *[a.ctypes.data_as(P) for a in t.arrays()[:5]], None,
                     t.node2point_indexstart.ctypes.data_as(P))
    t.node2point_index = np.zeros(ll, np.int64)
    L.hl_export(p.ctypes.data_as(P), n, topo.ctypes.data_as(P), M, perm.ctypes.data_as(P),
                *[a.ctypes.data_as(P) for a in t.arrays()])
    thr2 = topo[:M].view(np.float64)[:, 0]
    return t, thr2


def _cases():
    rng = np.random.default_rng(1)
    bad = rng.uniform(-1, 1, (500, 3))
    bad[::37, 1] = np.nan
    bad[5] = [0, 0, 0]
    bad[6] = [-0.0, 1.0, -1.0]
    bad[7] = [-1e-300, 1e-300, 0.5]
    bad[8] = [np.inf, -np.inf, 2.0]
    return {
        "sphere": synth.normalize(synth.sphere(4000, 1)).astype(np.float32).astype(np.float64),
        "torus": synth.normalize(synth.torus(3000, 2)),
        "duplicates": np.repeat(rng.uniform(-1, 1, (50, 3)), 40, axis=0),
        "outside_box": rng.uniform(-2, 2, (3000, 3)),
        "nan_and_edges": bad,
        "three": rng.uniform(-1, 1, (3, 3)),
        "one": rng.uniform(-1, 1, (1, 3)),
    }


@pytest.mark.parametrize("name", list(_cases()))
@pytest.mark.parametrize("S", [1, 16, 32, 64])
@pytest.mark.parametrize("md", [15, 6, 0])
def test_kernel_logic_replay_matches_oracle_tree(hl, name, S, md):
    p = _cases()[name]
    a = oracle.build_tree(p, md, S)
    b, thr2 = hl_tree(hl, p, md, S)
    for x, y in zip(a.arrays(), b.arrays()):
        assert np.array_equal(x, y)
    # opening thresholds: (2/theta * 2^-depth)^2 with theta = 0.5
    assert np.array_equal(thr2, (4.0 * a.half_w) ** 2)


def test_empty_cloud(hl):
    b, _ = hl_tree(hl, np.zeros((0, 3)), 15, 1)
    assert b.num_nodes == 0


def test_forest_replay_equals_per_cloud_trees(hl):
    """Clouds sorted together with the cloud id above the Morton bits form a
    forest whose per-cloud subtrees are exactly the standalone trees."""
    P = ct.c_void_p
    hl.hl_build_forest.argtypes = [P, P, ct.c_int64, ct.c_int, ct.c_int, ct.c_double, P, P, P]
    hl.hl_build_forest.restype = ct.c_int64
    rng = np.random.default_rng(4)
    clouds = [synth.normalize(synth.sphere(900, 1)), np.zeros((0, 3)), rng.uniform(-1, 1, (5, 3)),
              synth.normalize(synth.torus(1300, 2)), np.repeat(rng.uniform(-1, 1, (3, 3)), 30, 0),
              rng.uniform(-1, 1, (1, 3))]
    coff = np.cumsum([0] + [c.shape[0] for c in clouds]).astype(np.int64)
    pts = np.ascontiguousarray(np.concatenate(clouds), np.float64)
    for S in (1, 32):
        m = hl.hl_build_forest(pts.ctypes.data_as(P), coff.ctypes.data_as(P), len(clouds), 15, S, 0.5,
                               None, None, None)
        topo = np.zeros((m, 32), np.uint8)
        perm = np.zeros(pts.shape[0], np.int32)
        roots = np.zeros(len(clouds) + 1, np.int32)
        hl.hl_build_forest(pts.ctypes.data_as(P), coff.ctypes.data_as(P), len(clouds), 15, S, 0.5,
                           topo.ctypes.data_as(P), perm.ctypes.data_as(P), roots.ctypes.data_as(P))
        ints = topo.view(np.int32).reshape(m, 8)
        for c, cp in enumerate(clouds):
            r0, r1 = roots[c], roots[c + 1]
            p0 = coff[c]
            n_c = cp.shape[0]
            P2 = ct.c_void_p
            mm = np.zeros(1, np.int64)
            one = np.ascontiguousarray(cp, np.float64)
            hl.hl_build(one.ctypes.data_as(P2), n_c, 15, S, 0.5, mm.ctypes.data_as(P2), None, None)
            assert r1 - r0 == mm[0]
            if n_c == 0:
                continue
            t1 = np.zeros((mm[0], 32), np.uint8)
            perm1 = np.zeros(n_c, np.int32)
            hl.hl_build(one.ctypes.data_as(P2), n_c, 15, S, 0.5, mm.ctypes.data_as(P2),
                        t1.ctypes.data_as(P2), perm1.ctypes.data_as(P2))
            i1 = t1.view(np.int32).reshape(-1, 8)
            sub = ints[r0:r1]
            assert np.array_equal(sub[:, 2] - r0, i1[:, 2])          # escape
            assert np.array_equal(sub[:, 3] - p0, i1[:, 3])          # begin
            assert np.array_equal(sub[:, 4] - p0, i1[:, 4])          # end
            assert np.array_equal(sub[:, 5], i1[:, 5])               # depth | leaf
            par = sub[:, 6].copy()
            par[par >= 0] -= r0
            assert np.array_equal(par, i1[:, 6])                     # parent
            assert np.array_equal(topo[r0:r1].view(np.float64).reshape(-1, 4)[:, 0],
                                  t1.view(np.float64).reshape(-1, 4)[:, 0])  # thr2
            assert np.array_equal(perm[p0:p0 + n_c] - p0, perm1)


def hl_topo(L, p, md, S, theta=0.5):
    p = np.ascontiguousarray(p, np.float64)
    P = ct.c_void_p
    m = np.zeros(1, np.int64)
    L.hl_build(p.ctypes.data_as(P), p.shape[0], md, S, theta, m.ctypes.data_as(P), None, None)
    topo = np.zeros((max(int(m[0]), 1), 32), np.uint8)
    perm = np.zeros(max(p.shape[0], 1), np.int32)
    L.hl_build(p.ctypes.data_as(P), p.shape[0], md, S, theta, m.ctypes.data_as(P),
               topo.ctypes.data_as(P), perm.ctypes.data_as(P))
    return topo, int(m[0])


def hl_kids(L, topo, m):
    P = ct.c_void_p
    L.hl_kids.argtypes = [P, ct.c_int64, P, P, P]
    L.hl_kids.restype = None
    order, kids, nk = (np.zeros(max(m, 1), np.int32) for _ in range(3))
    L.hl_kids(topo.ctypes.data_as(P), m, order.ctypes.data_as(P), kids.ctypes.data_as(P),
              nk.ctypes.data_as(P))
    return order[:m], kids[:m], nk[:m]


@pytest.mark.parametrize("name", list(_cases()))
@pytest.mark.parametrize("S", [1, 32])
def test_child_positions_match_oracle_children(hl, name, S):
    """The upward pass reads a node's children at positions kids .. kids +
    nkids of the depth order: they must be exactly the oracle tree's children
    in octant order, and leaves have none."""
    p = _cases()[name]
    a = oracle.build_tree(p, 15, S)
    topo, m = hl_topo(hl, p, 15, S)
    assert m == a.num_nodes
    if m == 0:
        return
    order, kids, nk = hl_kids(hl, topo, m)
    depth = np.round(-np.log2(a.half_w)).astype(np.int64)
    assert np.all(np.diff(depth[order]) >= 0)
    for d in np.unique(depth):
        ids = order[depth[order] == d]
        assert np.all(np.diff(ids) > 0)     # ascending id within a depth
    for v in range(m):
        c = a.children[v][a.children[v] >= 0]
        if a.is_leaf[v]:
            assert kids[v] == -1 and nk[v] == 0
        else:
            assert 1 <= nk[v] <= 8
            assert np.array_equal(order[kids[v]:kids[v] + nk[v]], c), v


def test_child_positions_in_a_forest(hl):
    """In a forest (several clouds, one root each) every non-root node is
    listed once, under its own parent, children in ascending id."""
    P = ct.c_void_p
    hl.hl_build_forest.argtypes = [P, P, ct.c_int64, ct.c_int, ct.c_int, ct.c_double, P, P, P]
    hl.hl_build_forest.restype = ct.c_int64
    rng = np.random.default_rng(9)
    clouds = [synth.normalize(synth.sphere(700, 3)), rng.uniform(-1, 1, (4, 3)),
              synth.normalize(synth.torus(900, 5)), rng.uniform(-1, 1, (1, 3))]
    coff = np.cumsum([0] + [c.shape[0] for c in clouds]).astype(np.int64)
    pts = np.ascontiguousarray(np.concatenate(clouds), np.float64)
    for S in (1, 8):
        m = hl.hl_build_forest(pts.ctypes.data_as(P), coff.ctypes.data_as(P), len(clouds), 15, S,
                               0.5, None, None, None)
        topo = np.zeros((m, 32), np.uint8)
        perm = np.zeros(pts.shape[0], np.int32)
        roots = np.zeros(len(clouds) + 1, np.int32)
        hl.hl_build_forest(pts.ctypes.data_as(P), coff.ctypes.data_as(P), len(clouds), 15, S, 0.5,
                           topo.ctypes.data_as(P), perm.ctypes.data_as(P), roots.ctypes.data_as(P))
        parent = topo.view(np.int32).reshape(m, 8)[:, 6].copy()
        order, kids, nk = hl_kids(hl, topo, m)
        seen = np.zeros(m, np.int64)
        for v in range(m):
            if nk[v]:
                c = order[kids[v]:kids[v] + nk[v]]
                assert np.all(parent[c] == v) and np.all(np.diff(c) > 0)
                seen[c] += 1
        assert np.array_equal(seen, (parent >= 0).astype(np.int64))
        assert (parent < 0).sum() == len(clouds)
