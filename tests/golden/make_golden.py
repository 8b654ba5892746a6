"""Writes the golden fixtures in tests/golden/ (run once, in a container that
has the reference tree; the fixtures are committed, the reference is not).

* ref_core.npz -- vectors produced by the reference's OWN torch-free CPU core
  (CUDA-based/ext/wn_treecode/wn_treecode_cpu/*.cpp compiled into
  oracle/_ref/libcagr_ref.so by oracle/Makefile, driven by
  oracle/ref_harness.cpp): for three small seeded clouds and leaf sizes 1 and
  8, the 7 serialised tree arrays, the node moments of a vector and a scalar
  attribute, and multiply_by_A (CPU sign, continuous off/on), multiply_by_AT
  and multiply_by_G.
* bunny.npz -- the reference-held end-to-end fixture: the 5000 input points
  of CUDA-based/input/bunny.xyz and the normals main_CAGR.py produced for
  them, CUDA-based/results/bunny.xyz (columns 3..5).
* pgr.npz -- GaussRecon's PGR octree from the reference's OWN Octree::setTree
  (CUDA-based/ext/gaussrecon_src/*.cpp with its vendored ANN sources, compiled
  into oracle/_ref/libpgr_ref.so by oracle/Makefile, driven by
  oracle/pgr_ref_harness.cpp) for the float32 sample sets of PGR_CASES: the
  inputs, the counts, SHA-256 digests of the normalised samples, the leaves
  (depth, offsets in preorder) and the corners (gridDataVector order), and
  the full arrays for the smallest case.
* pgr_mesh.npz -- the reference's own marching cubes (Octree::initLeaf +
  GetMCIsoTriangles(iso, mesh, 0, 1, 0, 0) + writePolygon2, same library) on
  PGR octrees of MESH_CASES: digests of the float32 corner values it was fed,
  of its vertices, triangles and PLY bytes, and their counts.

Usage:  python tests/golden/make_golden.py [--ref /root/reference]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cagr_amd import synth  # noqa: E402

CLOUDS = {
    "sphere": lambda: synth.normalize(synth.sphere(600, 11)),
    "torus": lambda: synth.normalize(synth.torus(800, 12)),
    "clusters": lambda: np.repeat(np.random.default_rng(13).uniform(-0.9, 0.9, (40, 3)), 12, 0),
}
LEAVES = (1, 8)
WIDTH = 0.01


def ref_core_vectors():
    out = {}
    for name, gen in CLOUDS.items():
        p = gen().astype(np.float32).astype(np.float64)
        n = p.shape[0]
        rng = np.random.default_rng(len(name))
        mu = synth.random_normals(n, 100 + len(name))
        s = rng.normal(size=(n, 1))
        out[f"{name}/points"] = p
        out[f"{name}/mu"] = mu
        out[f"{name}/s"] = s
        for leaf in LEAVES:
            t = oracle.ref_build_tree(p, 15, leaf)
            key = f"{name}/leaf{leaf}"
            for field, arr in zip(("parent", "children", "is_leaf", "half_w", "num_points",
                                   "node2point_index", "node2point_indexstart"), t.arrays()):
                out[f"{key}/{field}"] = arr
            wmu = oracle.point_weights(mu)
            ws = oracle.point_weights(s)
            mm = oracle.ref_upward(t, p, mu, wmu)
            ms = oracle.ref_upward(t, p, s, ws)
            for field, arr in zip(("attrs", "reppoints", "weights"), mm):
                out[f"{key}/mu_{field}"] = arr
            for field, arr in zip(("attrs", "reppoints", "weights"), ms):
                out[f"{key}/s_{field}"] = arr
            out[f"{key}/A_cpu"] = oracle.ref_apply("A_cpu", t, p, mu, WIDTH, mm)
            out[f"{key}/A_cpu_cont"] = oracle.ref_apply("A_cpu", t, p, mu, WIDTH, mm, continuous=True)
            out[f"{key}/AT"] = oracle.ref_apply("AT", t, p, s, WIDTH, ms)
            out[f"{key}/G"] = oracle.ref_apply("G", t, p, mu, WIDTH, mm)
    out["width"] = np.array(WIDTH)
    return out


def pgr_inputs(bunny_points):
    r = np.random.default_rng
    plane = r(24).uniform(-1, 1, (1500, 3))
    plane[:, 2] = 0.25
    return {
        "bunny_d10_m1": (bunny_points, 10, 1),
        "bunny_d7_m3": (bunny_points, 7, 3),
        "sphere_d9_m1": (synth.sphere(4000, 21), 9, 1),
        "uniform_d8_m2": (r(22).uniform(-1, 1, (2000, 3)), 8, 2),
        "dup_d8_m1": (np.repeat(r(23).uniform(-0.5, 0.5, (100, 3)), 5, 0), 8, 1),
        "plane_d9_m1": (plane, 9, 1),
    }


def digest(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pgr_vectors(bunny_points):
    out = {}
    for name, (pts, d, m) in pgr_inputs(bunny_points).items():
        p = np.asarray(pts, np.float32)
        g = oracle.pgr_ref_build(p, d, m)
        if not name.startswith("bunny"):   # the bunny samples are bunny.npz's points
            out[f"{name}/points"] = p
        out[f"{name}/args"] = np.array([d, m, g.depth, len(g.leaves), len(g.corners)], np.int64)
        for field in ("samples", "leaves", "corners"):
            out[f"{name}/{field}_sha256"] = np.array(digest(getattr(g, field)))
        if len(g.leaves) < 12500:
            out[f"{name}/leaves"] = g.leaves
            out[f"{name}/corners"] = g.corners
    return out


def mesh_cases(bunny):
    pts = bunny["points"].astype(np.float32)
    return {
        # (points, max_depth, min_depth, field, smooth width)
        "bunny_d8_sphere": (pts, 8, 1, "sphere", 0.005),
        "bunny_d10_winding": (pts, 10, 1, "winding", 0.005),
        "bunny_d7_m2_winding": (pts, 7, 2, "winding", 0.005),
        "torus_d9_sphere": (synth.torus(3000, 51).astype(np.float32), 9, 1, "sphere", 0.004),
    }


def mesh_values(bunny, g, field):
    if field == "sphere":
        return oracle.sphere_field(g.corners), np.float32(0.0)
    return oracle.winding_field(g.samples, bunny["normals"], g.corners)


def pgr_mesh_vectors(bunny, tmpdir):
    out = {}
    for name, (pts, d, m, field, sw) in mesh_cases(bunny).items():
        if field == "winding" and len(pts) != len(bunny["normals"]):
            raise ValueError("winding fields need the bunny's normals")
        g = oracle.pgr_ref_build(pts, d, m)
        vals, iso = mesh_values(bunny, g, field)
        ply = os.path.join(tmpdir, name + ".ply")
        v, off, idx = oracle.pgr_ref_mesh(vals, iso, sw, ply)
        assert np.all(np.diff(off) == 3)
        if not name.startswith("bunny"):
            out[f"{name}/points"] = pts
        out[f"{name}/args"] = np.array([d, m], np.int64)
        out[f"{name}/field"] = np.array(field)
        out[f"{name}/iso_sw"] = np.array([iso, sw], np.float32)
        out[f"{name}/counts"] = np.array([len(v), len(off) - 1], np.int64)
        out[f"{name}/values_sha256"] = np.array(digest(vals))
        out[f"{name}/vertices_sha256"] = np.array(digest(v))
        out[f"{name}/triangles_sha256"] = np.array(digest(idx.reshape(-1, 3)))
        out[f"{name}/ply_sha256"] = np.array(hashlib_file(ply))
    return out


def hashlib_file(path):
    import hashlib
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def bunny(ref_root):
    base = os.path.join(ref_root, "CUDA-based")
    pin = np.loadtxt(os.path.join(base, "input", "bunny.xyz"))
    pout = np.loadtxt(os.path.join(base, "results", "bunny.xyz"))
    assert np.array_equal(pin[:, :3], pout[:, :3])
    return {"points": pin[:, :3], "normals": pout[:, 3:6]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    oracle.build()
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libcagr_ref.so is missing (make -C oracle needs the reference tree)")
    np.savez_compressed(os.path.join(HERE, "ref_core.npz"), **ref_core_vectors())
    bun = bunny(args.ref)
    np.savez_compressed(os.path.join(HERE, "bunny.npz"), **bun)
    if not oracle.pgr_ref_available():
        sys.exit("oracle/_ref/libpgr_ref.so is missing (make -C oracle needs the reference tree)")
    np.savez_compressed(os.path.join(HERE, "pgr.npz"), **pgr_vectors(bun["points"]))
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        np.savez_compressed(os.path.join(HERE, "pgr_mesh.npz"), **pgr_mesh_vectors(bun, tmp))
    for f in ("ref_core.npz", "bunny.npz", "pgr.npz", "pgr_mesh.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
