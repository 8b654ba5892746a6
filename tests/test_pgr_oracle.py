"""The PGR octree oracle (oracle/pgr_oracle.cpp) against the reference's own
Octree::setTree: the committed digests and arrays of tests/golden/pgr.npz
(written from oracle/_ref/libpgr_ref.so by tests/golden/make_golden.py), the
reference library itself where it was built, and the invariants an octree
query set must satisfy whatever its construction.  CPU only."""
import hashlib
import os

import numpy as np
import pytest

import oracle
from cagr_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PGR = np.load(os.path.join(GOLD, "pgr.npz"))
BUNNY = np.load(os.path.join(GOLD, "bunny.npz"))["points"].astype(np.float32)
CASES = sorted({k.split("/")[0] for k in PGR.files})


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def case_points(name):
    return BUNNY if name.startswith("bunny") else PGR[f"{name}/points"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_fixtures(name):
    d, m, depth, nl, nc = (int(v) for v in PGR[f"{name}/args"])
    g = oracle.pgr_build(case_points(name), d, m)
    assert (g.depth, len(g.leaves), len(g.corners)) == (depth, nl, nc)
    assert sha(g.samples) == str(PGR[f"{name}/samples_sha256"])
    assert sha(g.leaves) == str(PGR[f"{name}/leaves_sha256"])
    assert sha(g.corners) == str(PGR[f"{name}/corners_sha256"])
    if f"{name}/leaves" in PGR.files:
        assert np.array_equal(g.leaves, PGR[f"{name}/leaves"])
        assert np.array_equal(g.corners, PGR[f"{name}/corners"])


EXTRA = {
    "torus_d9_m1": (lambda: synth.torus(3000, 31), 9, 1),
    "ellipsoids_d10_m2": (lambda: synth.ellipsoids(6000, 32), 10, 2),
    "outliers_d8_m0": (lambda: np.concatenate([synth.sphere(2000, 33),
                                               np.random.default_rng(34).uniform(-1, 1, (100, 3))]),
                       8, 0),
    "two_points_d6_m1": (lambda: np.array([[0.1, 0.2, 0.3], [0.9, -0.4, 0.05]]), 6, 1),
    "line_d9_m1": (lambda: np.stack([np.linspace(-1, 1, 400)] * 3, 1), 9, 1),
    "min_equals_max_d5_m5": (lambda: synth.sphere(800, 35), 5, 5),
    "lattice_d7_m1": (lambda: np.stack(np.meshgrid(*[np.arange(9) / 8.0] * 3), -1).reshape(-1, 3),
                      7, 1),
}


@pytest.mark.skipif(not oracle.pgr_ref_available(),
                    reason="oracle/_ref/libpgr_ref.so not built (needs the reference tree)")
@pytest.mark.parametrize("name", sorted(EXTRA))
def test_oracle_matches_reference_build(name, tmp_path):
    gen, d, m = EXTRA[name]
    p = np.asarray(gen(), np.float32)
    a = oracle.pgr_ref_build(p, d, m, workdir=str(tmp_path))
    b = oracle.pgr_build(p, d, m)
    assert a.depth == b.depth
    assert np.array_equal(a.samples, b.samples)
    assert np.array_equal(a.leaves, b.leaves)
    assert np.array_equal(a.corners, b.corners)


def morton_left(leaves, H):
    d = leaves[:, 0].astype(np.int64)
    key = np.zeros(len(leaves), np.int64)
    for b in range(H):
        for a in range(3):
            bit = np.where(b < d, (leaves[:, 1 + a].astype(np.int64) >> np.maximum(d - 1 - b, 0)) & 1, 0)
            key |= bit << (3 * (H - 1 - b) + a)
    return key


@pytest.mark.parametrize("name", ["uniform_d8_m2", "dup_d8_m1", "plane_d9_m1"])
def test_octree_invariants(name):
    d, m, *_ = (int(v) for v in PGR[f"{name}/args"])
    g = oracle.pgr_build(case_points(name), d, m)
    lv = g.leaves.astype(np.int64)
    H = int(lv[:, 0].max())
    assert g.depth == H + 1
    # the leaves tile the unit cube: sum of 8^(H - d) = 8^H, offsets in range
    assert int((8 ** (H - lv[:, 0])).sum()) == 8 ** H
    assert np.all(lv[:, 1:] >= 0) and np.all(lv[:, 1:] < (1 << lv[:, :1]))
    # preorder = strictly increasing left-aligned Morton code (x lowest)
    assert np.all(np.diff(morton_left(lv, H)) > 0)
    # the samples lie inside the cube, the box fills 1/1.1 of it along its longest axis
    s = g.samples.astype(np.float64)
    assert s.min() > 0 and s.max() < 1
    ext = s.max(0) - s.min(0)
    assert abs(ext.max() - 1 / 1.1) < 1e-6
    # corners: unique, on the grid of depth H+1, each leaf's 8 corners present,
    # in order of first occurrence over (leaf preorder, corner 0..7)
    c = g.corners.astype(np.float64)
    ci = np.rint(c * (1 << H)).astype(np.int64)
    assert np.array_equal(ci, c * (1 << H))
    assert len(np.unique(ci, axis=0)) == len(ci)
    cell = (1 << (H - lv[:, 0]))[:, None]
    order = []
    seen = {}
    for r in range(len(lv)):
        for k in range(8):
            bits = np.array([(k >> a) & 1 for a in range(3)])
            key = tuple((lv[r, 1:] + bits) * cell[r])
            if key not in seen:
                seen[key] = len(order)
                order.append(key)
    assert np.array_equal(np.asarray(order), ci)
