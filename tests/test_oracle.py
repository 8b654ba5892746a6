"""Pins the oracle (oracle/cagr_oracle.cpp) before anything is checked against it.

* Golden vectors written by the reference's own compiled CPU core
  (tests/golden/ref_core.npz): tree topology, node moments, A (CPU sign),
  AT, G -- bit-for-bit.
* The live reference core (oracle/_ref) on fresh seeded clouds, when built.
* Brute-force direct sums (theta -> 0), the An = 0 reduction of the
  convection kernels to the plain Gauss kernels, the dense identity
  A2Tk(An) = (-A2k(-An))^T, the reference's dense AGR kernel
  (Cupy-based/utils_pgr.py:123-162) with its |c| terms removed, the sphere
  winding number, the exact-adjoint identity, and the reference-held CAGR
  result for bunny (tests/golden/bunny.npz).
"""
import os

import numpy as np
import pytest

import oracle
from cagr_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("parent", "children", "is_leaf", "half_w", "num_points", "node2point_index",
          "node2point_indexstart")


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLD, "ref_core.npz")))


def _clouds(g):
    return sorted({k.split("/")[0] for k in g if "/" in k})


def test_oracle_tree_matches_reference_golden(golden):
    for name in _clouds(golden):
        p = golden[f"{name}/points"]
        for leaf in (1, 8):
            t = oracle.build_tree(p, 15, leaf)
            for f, arr in zip(FIELDS, t.arrays()):
                assert np.array_equal(arr, golden[f"{name}/leaf{leaf}/{f}"]), (name, leaf, f)


def test_oracle_moments_and_plain_ops_match_reference_golden(golden):
    w = float(golden["width"])
    for name in _clouds(golden):
        p, mu, s = golden[f"{name}/points"], golden[f"{name}/mu"], golden[f"{name}/s"]
        for leaf in (1, 8):
            key = f"{name}/leaf{leaf}"
            t = oracle.build_tree(p, 15, leaf)
            mm = oracle.upward(t, p, mu)
            ms = oracle.upward(t, p, s)
            for f, a, b in zip(("attrs", "reppoints", "weights"), mm, ms):
                assert np.array_equal(a, golden[f"{key}/mu_{f}"]), (key, f)
                assert np.array_equal(b, golden[f"{key}/s_{f}"]), (key, f)
            assert np.array_equal(oracle.apply("A_cpu", t, p, mu, w, node_moments=mm),
                                  golden[f"{key}/A_cpu"])
            assert np.array_equal(oracle.apply("A_cpu", t, p, mu, w, node_moments=mm, continuous=True),
                                  golden[f"{key}/A_cpu_cont"])
            assert np.array_equal(oracle.apply("AT", t, p, s, w, node_moments=ms), golden[f"{key}/AT"])
            assert np.array_equal(oracle.apply("G", t, p, mu, w, node_moments=mm), golden[f"{key}/G"])
            # the CUDA sign convention is the exact negation
            assert np.array_equal(oracle.apply("A", t, p, mu, w, node_moments=mm), -golden[f"{key}/A_cpu"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("leaf", [1, 5, 32])
def test_oracle_matches_live_reference_core(leaf):
    rng = np.random.default_rng(leaf)
    clouds = [synth.normalize(synth.ellipsoids(2500, leaf)),
              np.concatenate([rng.uniform(-1.5, 1.5, (300, 3)), np.zeros((20, 3))]),  # outside box, duplicates
              ]
    for p in clouds:
        p = p.astype(np.float32).astype(np.float64)
        a = oracle.build_tree(p, 15, leaf)
        b = oracle.ref_build_tree(p, 15, leaf)
        bf = oracle.ref_build_tree(p.astype(np.float32), 15, leaf, f32=True)
        for x, y, z in zip(a.arrays(), b.arrays(), bf.arrays()):
            assert np.array_equal(x, y) and np.array_equal(x, z)
        mu = rng.normal(size=(p.shape[0], 3))
        w = oracle.point_weights(mu)
        m1, m2 = oracle.upward(a, p, mu, w), oracle.ref_upward(b, p, mu, w)
        for x, y in zip(m1, m2):
            assert np.array_equal(x, y)
        for cont in (False, True):
            assert np.array_equal(oracle.apply("A_cpu", a, p, mu, 0.02, node_moments=m1, continuous=cont),
                                  oracle.ref_apply("A_cpu", b, p, mu, 0.02, m2, continuous=cont))
        assert np.array_equal(oracle.apply("G", a, p, mu, 0.02, node_moments=m1),
                              oracle.ref_apply("G", b, p, mu, 0.02, m2))


def test_edge_cases_empty_and_single():
    t = oracle.build_tree(np.zeros((0, 3)), 15, 1)
    assert t.num_nodes == 0
    t = oracle.build_tree(np.array([[0.1, 0.2, 0.3]]), 15, 1)
    assert t.num_nodes == 1 and t.is_leaf[0] == 1
    # all points identical: recursion to depth 15, one leaf holding everything
    p = np.tile([[0.3, -0.2, 0.1]], (10, 1))
    t = oracle.build_tree(p, 15, 1)
    assert t.depth == 15 and t.num_nodes == 16 and t.num_points[-1] == 10


ALL_OPS = ["A", "A_cpu", "AT", "G", "A21", "A22", "A23", "A2T1", "A2T2", "A2T3", "A3"]


@pytest.mark.parametrize("op", ALL_OPS)
@pytest.mark.parametrize("cont", [False, True])
def test_theta_to_zero_is_direct_sum(op, cont):
    p = synth.normalize(synth.torus(700, 3))
    rng = np.random.default_rng(0)
    a = rng.normal(size=(p.shape[0], oracle.IN_DIM[op]))
    t = oracle.build_tree(p, 15, 4)
    w = rng.uniform(0.005, 0.03, p.shape[0])
    tc = oracle.apply(op, t, p, a, w, theta=1e-9, continuous=cont)
    d = oracle.direct(op, p, a, w, continuous=cont)
    scale = np.abs(d).max()
    assert np.abs(tc - d).max() <= 1e-12 * scale


@pytest.mark.parametrize("op", ALL_OPS)
def test_theta_controls_far_field_error(op):
    """The treecode error against the direct sum shrinks as theta shrinks."""
    p = synth.normalize(synth.sphere(1500, 4))
    a = np.random.default_rng(1).normal(size=(p.shape[0], oracle.IN_DIM[op]))
    t = oracle.build_tree(p, 15, 8)
    d = oracle.direct(op, p, a, 0.01)
    errs = [np.abs(oracle.apply(op, t, p, a, 0.01, theta=th) - d).max() for th in (0.8, 0.5, 0.25)]
    assert errs[0] > errs[1] > errs[2]


def test_an_zero_reduces_convection_kernels_to_gauss():
    p = synth.normalize(synth.ellipsoids(900, 5))
    rng = np.random.default_rng(2)
    mu = rng.normal(size=(p.shape[0], 3))
    s = rng.normal(size=(p.shape[0], 1))
    t = oracle.build_tree(p, 15, 6)
    for cont in (False, True):
        A = oracle.apply("A", t, p, mu, 0.02, continuous=cont)
        for k in ("A21", "A22", "A23"):
            assert np.array_equal(oracle.apply(k, t, p, mu, 0.02, an=0.0, continuous=cont), A)
        a3 = oracle.apply("A3", t, p, mu, 0.02, an=0.0, continuous=cont)
        assert np.allclose(a3, A, rtol=1e-14, atol=1e-14 * np.abs(A).max())
    AT = oracle.apply("AT", t, p, s, 0.02)
    for k in ("A2T1", "A2T2", "A2T3"):
        assert np.array_equal(oracle.apply(k, t, p, s, 0.02, an=0.0), AT)


def _dense(op, p, w, an, cont=False):
    """Dense matrix of the direct-sum op: rows = output (query, component),
    cols = input (source, component)."""
    n = p.shape[0]
    ind, outd = oracle.IN_DIM[op], oracle.OUT_DIM[op]
    M = np.zeros((n * outd, n * ind))
    for j in range(n):
        for c in range(ind):
            a = np.zeros((n, ind))
            a[j, c] = 1.0
            M[:, j * ind + c] = oracle.direct(op, p, a, w, an=an, continuous=cont).reshape(-1)
    return M


@pytest.mark.parametrize("k", [1, 2, 3])
def test_reference_transpose_identity(k):
    """A2Tk(An) equals (-A2k(-An))^T as dense matrices (the reference's
    "transpose" is the adjoint of the direction-reversed operator)."""
    p = synth.normalize(synth.sphere(40, 6))
    w = 0.05
    for an in (1.0, 0.7):
        T = _dense(f"A2T{k}", p, w, an)           # [(i,c), j]
        A = _dense(f"A2{k}", p, w, -an)           # [j, (i,c)]
        assert np.allclose(T, -A.T, rtol=1e-13, atol=1e-13 * np.abs(T).max())


def _agr_kernel_dense(x, y, width, c):
    """Restatement of the reference's dense AGR kernel get_A2
    (Cupy-based/utils_pgr.py:123-162): for query i, source j, component m,
    -(1/4pi) exp((c.d - |c| r_s)/2) (1/r_s) [d/r_s^2 + c/2 + |c| d/(2 r_s)],
    d = x_i - y_j, r_s = max(r, width); returns [Nq, 3, Ns]."""
    d = x[:, None, :] - y[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    rs = np.where(r > width, r, width)
    cn = np.linalg.norm(c)
    e = np.exp(0.5 * (d @ c - cn * rs)) / (4 * np.pi)
    inner = d / rs[..., None] ** 2 + 0.5 * c + 0.5 * cn * d / rs[..., None]
    return -(inner * (e / rs)[..., None]).transpose(0, 2, 1)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_convection_kernel_agrees_with_reference_agr_kernel(k):
    """Removing the AGR kernel's |c| terms (the treecode comments its Cn term
    out, cuda_kernels.cu:132) leaves -1/(4 pi) times the A2k kernel, so the
    convection sign and the c/2 term are pinned by a second reference file."""
    x = synth.normalize(synth.sphere(30, 7))
    y = synth.normalize(synth.sphere(25, 8)) * 0.8
    c = np.zeros(3)
    c[k] = 1.0
    G = _agr_kernel_dense(x, y, 1e-9, c)                       # [Nq, 3, Ns]
    d = x[:, None, :] - y[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    # undo the |c| terms: multiply back exp(|c| r / 2) and drop the |c| d / (2 r) part
    cn = 1.0
    G = G * np.exp(0.5 * cn * r)[:, None, :] + \
        np.exp(0.5 * (d @ c))[:, None, :] / (4 * np.pi) * (0.5 * cn * d / r[..., None] ** 2).transpose(0, 2, 1)
    mu = np.random.default_rng(k).normal(size=(y.shape[0], 3))
    agr = np.einsum("ims,sm->i", G, mu) * (-4 * np.pi)
    # oracle A2k with sources y, queries x (direct sum, no cutoff)
    ora = oracle.direct(f"A2{k + 1}", y, mu, 1e-9, queries=x)[:, 0]
    assert np.allclose(agr, ora, rtol=1e-12, atol=1e-12 * np.abs(ora).max())


def test_sphere_winding_number():
    """Unit sphere, outward normals * area: CUDA-sign A -> -4 pi inside, 0 outside."""
    n = 20000
    g = np.random.default_rng(9).normal(size=(n, 3))
    p = g / np.linalg.norm(g, axis=1, keepdims=True)
    mu = p * (4 * np.pi / n)
    q = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1], [0.0, 0.0, 2.0], [1.7, 0.4, -0.2]])
    val = oracle.direct("A", p, mu, 1e-9, queries=q)[:, 0]
    assert np.allclose(val[:2], -4 * np.pi, rtol=0.02)
    assert np.allclose(val[2:], 0.0, atol=0.05)
    t = oracle.build_tree(p, 15, 16)
    tv = oracle.apply("A", t, p, mu, 1e-9, queries=q)[:, 0]   # theta = 0.5 monopole far field
    assert np.allclose(tv[:2], -4 * np.pi, rtol=0.02)
    assert np.allclose(tv[2:], 0.0, atol=0.05)


@pytest.mark.parametrize("op", ["A", "A21", "A23", "A3", "AT", "A2T2", "G"])
def test_exact_adjoint_identity(op):
    """<A x, g> = <x, A* g> for the treecode with fixed centres, and A* equals
    the explicit transpose of the treecode's dense matrix."""
    p = synth.normalize(synth.ellipsoids(60, 10))
    n = p.shape[0]
    ind, outd = oracle.IN_DIM[op], oracle.OUT_DIM[op]
    t = oracle.build_tree(p, 15, 3)
    wfix = np.ones(n)
    rng = np.random.default_rng(3)
    _, rep, _ = oracle.upward(t, p, np.zeros((n, ind)), wfix)
    w = 0.02

    def fwd(x):
        na, _, _ = oracle.upward(t, p, x, wfix)
        return oracle.apply(op, t, p, x, w, node_moments=(na, rep), theta=0.7, continuous=True)

    M = np.zeros((n * outd, n * ind))
    for j in range(n * ind):
        e = np.zeros(n * ind)
        e[j] = 1.0
        M[:, j] = fwd(e.reshape(n, ind)).reshape(-1)
    g = rng.normal(size=(n, outd))
    adj = oracle.adjoint(op, t, p, g, w, rep, theta=0.7, continuous=True)
    assert np.allclose(adj.reshape(-1), M.T @ g.reshape(-1), rtol=1e-12,
                       atol=1e-12 * np.abs(adj).max())
    x = rng.normal(size=(n, ind))
    lhs = (fwd(x) * g).sum()
    rhs = (x * adj).sum()
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs))


def _direction_stats(a, b):
    cos = (a * b).sum(1) / np.linalg.norm(a, axis=1) / np.linalg.norm(b, axis=1)
    mag = np.linalg.norm(a, axis=1) / np.linalg.norm(b, axis=1)
    return np.median(cos), (cos > 0).mean(), np.median(mag)


def test_bunny_reference_result_reproduced():
    """The reference-held CAGR output (CUDA-based/results/bunny.xyz, normals of
    main_CAGR.py --width_config l0 on input/bunny.xyz) is reproduced by the
    oracle replay of the full schedule (40 SD/CG iterations + 8 WNNC steps)
    with convection strength An = 0: the same orientation for every point and
    the same magnitudes within a few percent (the reference ran in FP32 on an
    NVIDIA GPU).  With the source's An = 1 the replay does not reproduce it, so
    the file was written by a build with An = 0."""
    from oracle import solve_ref
    d = np.load(os.path.join(GOLD, "bunny.npz"))
    p = solve_ref.normalize(d["points"])
    normals, _ = solve_ref.cagr_schedule(p, *solve_ref.PRESETS["l0"], an=0.0)
    med_cos, agree, mag = _direction_stats(normals, d["normals"])
    assert med_cos > 0.9999 and agree == 1.0 and abs(mag - 1) < 0.02
    normals1, _ = solve_ref.cagr_schedule(p, *solve_ref.PRESETS["l0"], an=1.0)
    _, agree1, _ = _direction_stats(normals1, d["normals"])
    assert agree1 < 0.8
