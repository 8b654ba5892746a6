"""Marching cubes on the PGR octree, on the CPU: the per-leaf logic the HIP
kernels run (cagr_amd/csrc/mc_logic.h, compiled with g++ by
tests/host_logic/mc_replay.cpp and replayed serially) against the reference's
own GetMCIsoTriangles -- the committed digests of tests/golden/pgr_mesh.npz
(vertices, triangles and the PLY bytes of writePolygon2) and, where it was
built, oracle/_ref itself -- plus the PLY writer and output transform."""
import ctypes as ct
import hashlib
import os
import subprocess

import numpy as np
import pytest

import oracle
from cagr_amd import synth
from cagr_amd.pcio import save_ply_mesh
from cagr_amd.pgr import output_transform

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MESH = np.load(os.path.join(GOLD, "pgr_mesh.npz"))
BUNNY = np.load(os.path.join(GOLD, "bunny.npz"))
CASES = sorted({k.split("/")[0] for k in MESH.files})


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("mc") / "mc_replay.so")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-fsanitize=undefined", "-fno-sanitize-recover=all", "-D_GLIBCXX_ASSERTIONS",
                    "-o", so,
                    os.path.join(HERE, "host_logic", "mc_replay.cpp")], check=True)
    L = ct.CDLL(so)
    L.hl_mesh.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_float, ct.c_float,
                          ct.c_void_p, ct.c_void_p]
    L.hl_mesh_fetch.argtypes = [ct.c_void_p, ct.c_void_p]
    return L


def run_replay(L, g, vals, iso, sw):
    lv = np.ascontiguousarray(g.leaves)
    v = np.ascontiguousarray(vals, np.float32)
    nv, nt = ct.c_int64(), ct.c_int64()
    assert L.hl_mesh(lv.ctypes.data, len(lv), g.depth, v.ctypes.data, float(iso), float(sw),
                     ct.byref(nv), ct.byref(nt)) == 0
    P = np.zeros((nv.value, 3), np.float32)
    T = np.zeros((nt.value, 3), np.int32)
    L.hl_mesh_fetch(P.ctypes.data, T.ctypes.data)
    return P, T


def case_inputs(name):
    d, m = (int(x) for x in MESH[f"{name}/args"])
    pts = BUNNY["points"].astype(np.float32) if name.startswith("bunny") else MESH[f"{name}/points"]
    g = oracle.pgr_build(pts, d, m)
    field = str(MESH[f"{name}/field"])
    if field == "sphere":
        vals, iso = oracle.sphere_field(g.corners), np.float32(0.0)
    else:
        vals, iso = oracle.winding_field(g.samples, BUNNY["normals"], g.corners)
    iso_f, sw = (np.float32(x) for x in MESH[f"{name}/iso_sw"])
    assert iso == iso_f and sha(vals) == str(MESH[f"{name}/values_sha256"])
    return pts, g, vals, iso, sw


def box_scale(pts):
    p = np.asarray(pts, np.float32)
    mn, mx = p.min(0), p.max(0)
    return np.concatenate([mn, mx]), np.float32((mx - mn).max()) * np.float32(1.1)


@pytest.mark.parametrize("name", CASES)
def test_mesh_replay_matches_reference_fixtures(replay, name, tmp_path):
    pts, g, vals, iso, sw = case_inputs(name)
    P, T = run_replay(replay, g, vals, iso, sw)
    assert [len(P), len(T)] == MESH[f"{name}/counts"].tolist()
    assert sha(P) == str(MESH[f"{name}/vertices_sha256"])
    assert sha(T) == str(MESH[f"{name}/triangles_sha256"])
    box, scale = box_scale(pts)
    path = str(tmp_path / "mesh.ply")
    save_ply_mesh(path, output_transform(P, box, scale), T)
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == str(MESH[f"{name}/ply_sha256"])


EXTRA = {
    "ellipsoids_d9_sphere": (lambda: synth.ellipsoids(4000, 61), 9, 1, 0.3),
    "lattice_d7_sphere": (lambda: np.stack(np.meshgrid(*[np.arange(9) / 8.0] * 3), -1).reshape(-1, 3),
                          7, 1, 0.27),
    "uniform_d8_m3_sphere": (lambda: np.random.default_rng(62).uniform(-1, 1, (2500, 3)), 8, 3, 0.2),
}


@pytest.mark.skipif(not oracle.pgr_ref_available(),
                    reason="oracle/_ref/libpgr_ref.so not built (needs the reference tree)")
@pytest.mark.parametrize("name", sorted(EXTRA))
def test_mesh_replay_matches_reference_build(replay, name, tmp_path):
    gen, d, m, r = EXTRA[name]
    pts = np.asarray(gen(), np.float32)
    g = oracle.pgr_ref_build(pts, d, m, workdir=str(tmp_path))
    vals = oracle.sphere_field(g.corners, radius=r)
    ply = str(tmp_path / "ref.ply")
    v, off, idx = oracle.pgr_ref_mesh(vals, 0.0, 0.004, ply)
    P, T = run_replay(replay, g, vals, 0.0, 0.004)
    assert np.all(np.diff(off) == 3)
    assert np.array_equal(P, v) and np.array_equal(T, idx.reshape(-1, 3))
    box, scale = box_scale(pts)
    mine = str(tmp_path / "mine.ply")
    save_ply_mesh(mine, output_transform(P, box, scale), T)
    assert open(mine, "rb").read() == open(ply, "rb").read()


def test_mesh_is_a_closed_surface_for_an_interior_sphere(replay):
    """A sphere well inside the cube gives a closed, consistently oriented
    triangle mesh: every edge is shared by exactly two triangles, once in each
    direction, and the Euler characteristic V - E + F is 2."""
    pts = BUNNY["points"].astype(np.float32)
    g = oracle.pgr_build(pts, 8, 1)
    vals = oracle.sphere_field(g.corners, radius=0.3)
    P, T = run_replay(replay, g, vals, 0.0, 0.005)
    e = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    directed = {tuple(x) for x in e}
    assert len(directed) == len(e)
    assert all((b, a) in directed for a, b in directed)
    n_edges = len(e) // 2
    assert len(P) - n_edges + len(T) == 2
    r = np.linalg.norm(P.astype(np.float64) - 0.5, axis=1)
    assert np.abs(r - 0.3).max() < 0.02


def test_workspace_index_check_aborts_out_of_range(tmp_path):
    """The replay's workspace index check is live: one index past an array's
    length aborts the program (the kernels compile the same accessors with the
    check off)."""
    src = tmp_path / "oob.cpp"
    src.write_text(
        '#define CAGR_HD\n#define CAGR_MC_STRIDE 1\n#include <cstdio>\n#include <cstdlib>\n'
        '#define CAGR_MC_CHECK(c) ((c) ? (void)0 : (std::fputs("oob\\n", stderr), std::abort()))\n'
        '#include <cmath>\nusing std::sqrt;\n'
        f'#include "{os.path.join(HERE, "..", "cagr_amd", "csrc", "mc_logic.h")}"\n'
        'using namespace cagr::mc;\n'
        'alignas(8) static unsigned char blk[WorkLayout::bytes];\n'
        'int main(int argc, char**) {\n'
        '    Work w(blk, 0);\n'
        '    w.ea[kMaxEdges - 1] = 1;\n'
        '    w.own[11].key = 2;\n'
        '    if (argc > 1) w.ea[kMaxEdges] = 3;\n'
        '    return 0;\n}\n')
    exe = str(tmp_path / "oob")
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", exe, str(src)], check=True)
    assert subprocess.run([exe]).returncode == 0
    r = subprocess.run([exe, "x"], capture_output=True, text=True)
    assert r.returncode != 0 and "oob" in r.stderr
