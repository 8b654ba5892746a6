"""Point-cloud input/output of cagr_amd.pcio against main_CAGR.py's host code
(CUDA-based/main_CAGR.py:27-45, 188-192): the normalization's closed form,
every input format read back exactly, the output byte-identical to
numpy.savetxt of the same rows.  PLY/OBJ files are written here by hand
(trimesh, the reference's reader, is not installed), in the three PLY
encodings, with extra properties and elements around the vertices."""
import os

import numpy as np
import pytest

from cagr_amd import pcio, synth


def _pts(n=37, seed=3):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 3)) * [3.0, 0.5, 1.5] + [10.0, -2.0, 0.25]


def test_normalize_closed_form():
    p = _pts()
    q, c, s = pcio.normalize(p)
    lo, hi = p.min(0), p.max(0)
    assert np.array_equal(c, (lo + hi) / 2.0)
    assert s == 2 / ((hi - lo).max() * 1.1)
    assert np.array_equal(q, (p - c) * s)
    # the normalized cloud is centred, its longest side is 2 / 1.1
    assert np.allclose((q.min(0) + q.max(0)) / 2, 0.0, atol=1e-15)
    assert np.isclose((q.max(0) - q.min(0)).max(), 2 / 1.1, rtol=1e-15)
    assert np.array_equal(q, synth.normalize(p))


def test_xyz_npy_obj_round_trip(tmp_path):
    p = _pts()
    n = _pts(seed=4)
    f = str(tmp_path / "a.xyz")
    np.savetxt(f, np.concatenate([p, n], 1))
    a, an = pcio.load_points(f)
    assert np.array_equal(a, p) and np.array_equal(an, n)
    np.savetxt(f, p)
    a, an = pcio.load_points(f)
    assert np.array_equal(a, p) and an is None
    f = str(tmp_path / "a.npy")
    np.save(f, np.concatenate([p, n], 1))
    a, _ = pcio.load_points(f)
    assert np.array_equal(a, p)
    f = str(tmp_path / "a.obj")
    with open(f, "w") as fh:
        fh.write("# comment\no thing\n")
        for v in p:
            fh.write("v %.17g %.17g %.17g 1.0\n" % tuple(v))
        fh.write("vn 0 0 1\nvt 0.5 0.5\nf 1 2 3\n")
    a, _ = pcio.load_points(f)
    assert np.array_equal(a, p)


def _ply_header(fmt, n, extra_before=0):
    h = "ply\nformat %s 1.0\ncomment made by a test\n" % fmt
    if extra_before:
        h += "element camera %d\nproperty float view_x\nproperty uchar id\n" % extra_before
    h += ("element vertex %d\nproperty double x\nproperty float red_ish\nproperty double y\n"
          "property double z\nproperty float nx\nproperty float ny\nproperty float nz\n"
          "property uchar alpha\n" % n)
    h += "element face 1\nproperty list uchar int vertex_indices\nend_header\n"
    return h


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_vertices(tmp_path, fmt):
    p = _pts(23)
    nrm = np.float32(_pts(23, seed=8))
    f = str(tmp_path / "a.ply")
    before = 0 if fmt == "ascii" else 2
    with open(f, "wb") as fh:
        fh.write(_ply_header(fmt, len(p), before).encode())
        if fmt == "ascii":
            for v, nv in zip(p, nrm):
                fh.write(("%.17g 0.5 %.17g %.17g %.9g %.9g %.9g 7\n"
                          % (v[0], v[1], v[2], nv[0], nv[1], nv[2])).encode())
            fh.write(b"3 0 1 2\n")
        else:
            e = "<" if fmt.endswith("little_endian") else ">"
            cam = np.zeros(before, dtype=[("view_x", e + "f4"), ("id", "u1")])
            fh.write(cam.tobytes())
            rec = np.zeros(len(p), dtype=[("x", e + "f8"), ("r", e + "f4"), ("y", e + "f8"),
                                          ("z", e + "f8"), ("nx", e + "f4"), ("ny", e + "f4"),
                                          ("nz", e + "f4"), ("a", "u1")])
            rec["x"], rec["y"], rec["z"] = p[:, 0], p[:, 1], p[:, 2]
            rec["nx"], rec["ny"], rec["nz"] = nrm[:, 0], nrm[:, 1], nrm[:, 2]
            fh.write(rec.tobytes())
            fh.write(np.array([3], "u1").tobytes() + np.array([0, 1, 2], e + "i4").tobytes())
    a, an = pcio.load_points(f)
    assert np.array_equal(a, p)
    assert np.array_equal(an, nrm.astype(np.float64))


def test_bad_inputs(tmp_path):
    with pytest.raises(NotImplementedError):
        pcio.load_points(str(tmp_path / "a.txt"))
    f = str(tmp_path / "b.ply")
    with open(f, "w") as fh:
        fh.write("ply\nformat ascii 1.0\nelement face 0\nproperty list uchar int v\nend_header\n")
    with pytest.raises(ValueError):
        pcio.load_points(f)
    with open(f, "wb") as fh:
        fh.write(_ply_header("binary_little_endian", 5).encode() + b"\x00" * 10)
    with pytest.raises(ValueError):
        pcio.load_points(f)


def test_save_xyz_is_numpy_savetxt(tmp_path):
    p, n = _pts(), _pts(seed=5)
    f1, f2 = str(tmp_path / "1.xyz"), str(tmp_path / "2.xyz")
    pcio.save_xyz(f1, p, n)
    np.savetxt(f2, np.concatenate([p, n], -1))
    with open(f1, "rb") as a, open(f2, "rb") as b:
        assert a.read() == b.read()
    a, an = pcio.load_points(f1)
    assert np.array_equal(a, p) and np.array_equal(an, n)


def test_orient_rejects_unknown_width_config(tmp_path):
    with pytest.raises(ValueError):
        pcio.orient(str(tmp_path / "x.xyz"), "l9")
    with pytest.raises(ValueError):
        pcio.orient(str(tmp_path / "x.xyz"), "custom", wsmin=0.01)
