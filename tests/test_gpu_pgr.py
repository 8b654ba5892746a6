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
from cagr_amd.pgr import PGROctree, gauss_recon, gauss_recon_values
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


# ---------------------------------------------------------------- marching cubes
MESH = np.load(os.path.join(GOLD, "pgr_mesh.npz"))
MESH_CASES = sorted({k.split("/")[0] for k in MESH.files})


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", MESH_CASES)
def test_pgr_mesh_matches_reference_fixtures(name, tmp_path):
    """Vertices, triangles and the PLY bytes identical to the reference's own
    GetMCIsoTriangles + writePolygon2 on the same corner values."""
    from cagr_amd.pcio import save_ply_mesh
    d, m = (int(x) for x in MESH[f"{name}/args"])
    pts = BUNNY["points"].astype(np.float32) if name.startswith("bunny") else MESH[f"{name}/points"]
    say("pgr mesh", name)
    o = oracle.pgr_build(pts, d, m)
    if str(MESH[f"{name}/field"]) == "sphere":
        vals, iso = oracle.sphere_field(o.corners), np.float32(0.0)
    else:
        vals, iso = oracle.winding_field(o.samples, BUNNY["normals"], o.corners)
    assert _sha(vals) == str(MESH[f"{name}/values_sha256"])
    sw = float(MESH[f"{name}/iso_sw"][1])
    g = PGROctree(torch.from_numpy(pts).to(DEV), d, m)
    v, t = g.mesh(torch.from_numpy(vals).to(DEV), float(iso), sw)
    assert [v.shape[0], t.shape[0]] == MESH[f"{name}/counts"].tolist()
    assert _sha(v.cpu().numpy()) == str(MESH[f"{name}/vertices_sha256"])
    assert _sha(t.cpu().numpy()) == str(MESH[f"{name}/triangles_sha256"])
    path = str(tmp_path / "m.ply")
    save_ply_mesh(path, g.output_vertices(v), t.cpu().numpy())
    assert _sha(np.frombuffer(open(path, "rb").read(), np.uint8)) == str(MESH[f"{name}/ply_sha256"])
    # a second mesh on the same octree replaces the first and is the same
    v2, t2 = g.mesh(torch.from_numpy(vals).to(DEV), float(iso), sw)
    assert torch.equal(v, v2) and torch.equal(t, t2)


def test_pgr_mesh_empty_when_no_corner_crosses_the_iso_value(tmp_path):
    """An iso-value above every corner value: no edge root, no triangle -- an
    empty mesh and a valid PLY with zero elements, as the reference's
    GetMCIsoTriangles + writePolygon2 leave; the octree still meshes after."""
    from cagr_amd.pcio import save_ply_mesh
    pts = BUNNY["points"].astype(np.float32)
    say("pgr mesh empty")
    g = PGROctree(torch.from_numpy(pts).to(DEV), 8, 1)
    vals = oracle.sphere_field(g.corners.cpu().numpy())
    v, t = g.mesh(torch.from_numpy(vals).to(DEV), float(vals.max()) + 1.0, 0.005)
    assert v.shape == (0, 3) and t.shape == (0, 3)
    path = str(tmp_path / "empty.ply")
    save_ply_mesh(path, g.output_vertices(v), t.cpu().numpy())
    data = open(path, "rb").read()
    assert b"element vertex 0\n" in data and b"element face 0\n" in data
    assert data.endswith(b"end_header\n")
    v2, t2 = g.mesh(torch.from_numpy(vals).to(DEV), 0.0, 0.005)
    assert v2.shape[0] > 0 and t2.shape[0] > 0


def test_gauss_recon_end_to_end(tmp_path):
    """main_GaussRecon_cuda.cu as one call on the bunny's oriented samples: the
    FP64 field differs from the oracle's only in summation order, so the mesh
    matches the reference's on the oracle's field in size and lies within a
    float32 rounding of it."""
    from cagr_amd.pgr import gauss_recon
    say("gauss recon end to end")
    pts = torch.from_numpy(BUNNY["points"].astype(np.float32)).to(DEV)
    nml = torch.from_numpy(BUNNY["normals"]).to(DEV)
    path = str(tmp_path / "bunny.ply")
    v, t, iso = gauss_recon(pts, nml, path, max_depth=10, precision="fp64")
    nv, nt = MESH["bunny_d10_winding/counts"].tolist()
    assert abs(len(v) - nv) <= 2 and abs(len(t) - nt) <= 4
    assert abs(iso - float(MESH["bunny_d10_winding/iso_sw"][0])) <= 1e-6
    data = open(path, "rb").read()
    assert data.startswith(b"ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % len(v))
    lo, hi = BUNNY["points"].min(0), BUNNY["points"].max(0)
    span = (hi - lo).max()
    assert np.all(v >= lo - 0.1 * span) and np.all(v <= hi + 0.1 * span)
    # FP32c: the same surface up to the operator's tolerance
    v32, t32, _ = gauss_recon(pts, nml, None, max_depth=10, precision="fp32c")
    assert abs(len(v32) - nv) <= 0.01 * nv


def test_reconstruct_from_an_oriented_point_file(tmp_path):
    """pcio.reconstruct = main_GaussReconCUDA -i file -o mesh.ply: the same mesh
    and PLY bytes as gauss_recon on the float32 samples and normals."""
    from cagr_amd import pcio
    from cagr_amd.pgr import gauss_recon
    say("reconstruct from file")
    pts = BUNNY["points"].astype(np.float32)
    nml = BUNNY["normals"].astype(np.float32)
    src = str(tmp_path / "bunny.xyz")
    np.savetxt(src, np.hstack([pts, nml]), fmt="%.9g")
    a = str(tmp_path / "a.ply")
    b = str(tmp_path / "b.ply")
    v1, t1, i1 = pcio.reconstruct(src, a, max_depth=8, precision="fp64")
    v2, t2, i2 = gauss_recon(torch.from_numpy(pts).to(DEV), torch.from_numpy(nml).to(DEV), b,
                             max_depth=8, precision="fp64")
    assert i1 == i2 and np.array_equal(v1, v2) and np.array_equal(t1, t2)
    assert open(a, "rb").read() == open(b, "rb").read()


@pytest.mark.parametrize("name", ["line", "lattice", "outliers", "sparse"])
def test_pgr_octree_degenerate_and_sparse_inputs(name):
    """Collinear samples, exact cell-boundary ties, outliers and sparse
    samples (whose depth search climbs to the shallow depths that the build
    defers to its second pass) match the oracle exactly."""
    rng = np.random.default_rng(71)
    p = {"line": np.stack([np.linspace(-1, 1, 400)] * 3, 1),
         "lattice": np.stack(np.meshgrid(*[np.arange(9) / 8.0] * 3), -1).reshape(-1, 3),
         "outliers": np.concatenate([synth.sphere(3000, 72), rng.uniform(-1, 1, (150, 3))]),
         "sparse": rng.uniform(-1, 1, (40, 3))}[name].astype(np.float32)
    say("pgr degenerate", name)
    for d, m in ((9, 1), (6, 0)):
        o = oracle.pgr_build(p, d, m)
        same(PGROctree(torch.from_numpy(p).to(DEV), d, m), o)



def test_gauss_recon_cpu_driver_field():
    """main_GaussRecon_cpu.cpp:303-359: the field is the CPU-sign plain
    operator multiply_by_A (continuous kernel) at the corners and samples."""
    pts = BUNNY["points"].astype(np.float32)
    nml = BUNNY["normals"]
    say("gauss recon cpu-driver field")
    o = oracle.pgr_build(pts, 8, 1)
    p64 = (2 * o.samples - 1).astype(np.float64)
    q64 = (2 * o.corners - 1).astype(np.float64)
    mu = sample_attributes(torch.from_numpy(nml).to(DEV).float()).double().cpu().numpy()
    ot = oracle.build_tree(p64, 15, 1)
    ref_q = oracle.apply("A_cpu", ot, p64, mu, 0.01, queries=q64, continuous=True)[:, 0]
    ref_s = oracle.apply("A_cpu", ot, p64, mu, 0.01, continuous=True)[:, 0]
    ref_iso = -np.partition(ref_s, len(ref_s) // 2)[len(ref_s) // 2]
    t_pts = torch.from_numpy(pts).to(DEV)
    t_nml = torch.from_numpy(nml).to(DEV)
    for prec, tol in (("fp64", 1e-12), ("fp32c", 2e-6)):
        oct_, vals, iso, sw = gauss_recon_values(t_pts, t_nml, max_depth=8, precision=prec,
                                                 op="A_cpu")
        v = vals.double().cpu().numpy()
        scale = np.abs(ref_q).max()
        assert np.abs(v + ref_q).max() <= tol * scale + 1e-7 * scale, prec
        assert abs(iso - ref_iso) <= 1e-6 * np.abs(ref_s).max(), prec
    # the whole driver: a closed mesh from the CPU-driver field
    verts, tris, iso = gauss_recon(t_pts, t_nml, None, max_depth=8, precision="fp64", op="A_cpu")
    assert abs(iso - ref_iso) <= 1e-6 * np.abs(ref_s).max()
    assert len(verts) > 100 and len(tris) > 100
    assert tris.min() >= 0 and tris.max() < len(verts)
    with pytest.raises(ValueError):
        gauss_recon_values(t_pts, t_nml, max_depth=8, op="G")
