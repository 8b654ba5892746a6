"""GPU parity of the PGR octree builder (cagr_pgr_build) and of the
GaussRecon values on its corners.

Bars:
* samples, leaves, corners and depth: identical to the oracle (the octree is
  integer structure and the coordinates are exact dyadic floats; the float
  sums behind it are taken in the reference's order);
* winding numbers at the corners: the A3 bars of the operator tests (FP64
  1e-12, FP32-compensated 2e-6, normwise against the FP64 oracle).
"""
import os

import numpy as np
import pytest
import torch

import oracle
from cagr_amd import _lib, synth
from cagr_amd.pgr import PGROctree, gauss_recon_values
from cagr_amd.query import sample_attributes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PGR = np.load(os.path.join(GOLD, "pgr.npz"))
BUNNY = np.load(os.path.join(GOLD, "bunny.npz"))


def say(*a):
    print("[case]", *a, flush=True)


def case_points(name):
    if name.startswith("bunny"):
        return BUNNY["points"].astype(np.float32)
    return PGR[f"{name}/points"]


def same(g, o):
    assert g.depth == o.depth
    assert torch.equal(g.samples.cpu(), torch.from_numpy(o.samples))
    assert torch.equal(g.leaves.cpu(), torch.from_numpy(o.leaves))
    assert torch.equal(g.corners.cpu(), torch.from_numpy(o.corners))


CASES = sorted({k.split("/")[0] for k in PGR.files})


@pytest.mark.parametrize("name", CASES)
def test_pgr_octree_matches_oracle(name):
    d, m, *_ = (int(v) for v in PGR[f"{name}/args"])
    p = case_points(name)
    say("pgr", name, p.shape, d, m)
    same(PGROctree(torch.from_numpy(p).to(DEV), d, m), oracle.pgr_build(p, d, m))


@pytest.mark.parametrize("n,d,m,seed", [(2, 6, 1, 0), (50000, 10, 1, 1), (120000, 11, 2, 2),
                                        (30000, 13, 1, 3), (20000, 3, 0, 4)])
def test_pgr_octree_sizes(n, d, m, seed):
    p = synth.ellipsoids(n, 40 + seed).astype(np.float32) if n > 2 else \
        np.array([[0.1, 0.2, 0.3], [0.9, -0.4, 0.05]], np.float32)
    say("pgr sizes", n, d, m, seed)
    same(PGROctree(torch.from_numpy(p).to(DEV), d, m), oracle.pgr_build(p, d, m))


def test_pgr_octree_is_deterministic_and_rejects_bad_input():
    p = torch.from_numpy(synth.sphere(5000, 41).astype(np.float32)).to(DEV)
    say("pgr repeat")
    a, b = PGROctree(p, 9, 1), PGROctree(p, 9, 1)
    assert torch.equal(a.corners, b.corners) and torch.equal(a.leaves, b.leaves)
    with pytest.raises(_lib.CagrError, match="box"):
        PGROctree(torch.ones((10, 3), device=DEV), 8, 1)
    with pytest.raises(_lib.CagrError, match="max_depth"):
        PGROctree(p, 14, 1)
    with pytest.raises(TypeError):
        PGROctree(p.double(), 8, 1)


def test_gauss_recon_values_on_the_pgr_corners():
    pts = BUNNY["points"].astype(np.float32)
    nml = BUNNY["normals"]
    say("gauss recon bunny")
    o = oracle.pgr_build(pts, 8, 1)
    p64 = (2 * o.samples - 1).astype(np.float64)
    q64 = (2 * o.corners - 1).astype(np.float64)
    # the attributes as the product forms them (float32 normals, unit * 1e-5)
    mu = sample_attributes(torch.from_numpy(nml).to(DEV).float()).double().cpu().numpy()
    ot = oracle.build_tree(p64, 15, 1)
    ref_q = oracle.apply("A3", ot, p64, mu, 0.01, queries=q64, continuous=True)[:, 0]
    ref_s = oracle.apply("A3", ot, p64, mu, 0.01, continuous=True)[:, 0]
    ref_iso = -np.partition(ref_s, len(ref_s) // 2)[len(ref_s) // 2]
    for prec, tol in (("fp64", 1e-12), ("fp32c", 2e-6)):
        t_pts = torch.from_numpy(pts).to(DEV)
        t_nml = torch.from_numpy(nml).to(DEV)
        oct_, vals, iso, sw = gauss_recon_values(t_pts, t_nml, max_depth=8, precision=prec)
        assert torch.equal(oct_.corners.cpu(), torch.from_numpy(o.corners))
        v = vals.double().cpu().numpy()
        assert np.abs(v + ref_q).max() <= tol * np.abs(ref_q).max() + 1e-7 * np.abs(ref_q).max(), prec
        assert abs(iso - ref_iso) <= 1e-6 * np.abs(ref_s).max(), prec
        assert sw == 0.005
        # the surface separates the samples' inside from the far corners:
        # the isovalue lies strictly between the values far outside and deep inside
        assert v.min() < iso < v.max()
