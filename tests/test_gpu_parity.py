"""GPU parity of the cagr_amd HIP library against the oracle and the golden
fixtures written by the reference's own CPU core.

Bars (each with its reason):
* tree topology, the sorted-point order and per-query interaction counts:
  identical (integer work);
* FP64 node moments: bit-identical (same operations in the same order, no
  contraction);
* FP64 operators: max |gpu - oracle| <= 1e-12 * max |oracle| (same terms and
  decisions, different summation order and 1/r evaluation: observed ~2e-15);
* FP32-compensated operators: max |gpu - oracle| <= 2e-6 * max |oracle|
  (FP32 kernel sums with compensated accumulation: observed <= 6.3e-7);
* results returned in float32 carry one more float32 rounding (<= 1e-7);
* FP32c on FP64 points that are not FP32 values, in pairs 5e-4 apart with
  widths on the pairs' distances: 1e-3 (the float copies' own rounding; a
  wrong cutoff decision shows as 0.01-1).
"""
import os

import numpy as np
import pytest
import torch

import oracle
from cagr_amd import WindingNumberTreecode, synth

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL64 = 1e-12
TOL32C = 2e-6
TOLF64PTS = 1e-3   # FP32c on FP64 points in close pairs (see the test)


def say(*a):
    print("[case]", *a, flush=True)


def normwise(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def t(x, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV, dtype)


CLOUDS = {
    "sphere3k": lambda: synth.normalize(synth.sphere(3000, 1)),
    "torus4k_outliers": lambda: synth.normalize(synth.torus(4000, 2)),
    "ellipsoids5k": lambda: synth.normalize(synth.ellipsoids(5000, 3)),
}


@pytest.fixture(scope="module", params=list(CLOUDS))
def cloud(request):
    p = CLOUDS[request.param]().astype(np.float32).astype(np.float64)
    return request.param, p


@pytest.mark.parametrize("leaf", [1, 32])
def test_tree_topology_matches_oracle(cloud, leaf):
    name, p = cloud
    say("tree", name, leaf)
    for dt in (torch.float32, torch.float64):
        wt = WindingNumberTreecode(t(p, dt), leaf_size=leaf)
        ot = oracle.build_tree(p, 15, leaf)
        for a, b in zip(wt.reference_layout(), ot.arrays()):
            assert np.array_equal(a, b)
        info = wt.info()
        assert info["num_nodes"] == ot.num_nodes and info["depth"] == ot.depth


def test_tree_and_moments_match_reference_golden():
    g = dict(np.load(os.path.join(GOLD, "ref_core.npz")))
    for name in ("sphere", "torus", "clusters"):
        p = g[f"{name}/points"]
        for leaf in (1, 8):
            say("golden", name, leaf)
            key = f"{name}/leaf{leaf}"
            wt = WindingNumberTreecode(t(p), leaf_size=leaf)
            layout = wt.reference_layout()
            for f, arr in zip(("parent", "children", "is_leaf", "half_w", "num_points",
                               "node2point_index", "node2point_indexstart"), layout):
                assert np.array_equal(arr, g[f"{key}/{f}"]), (key, f)
            for attr, a in (("mu", g[f"{name}/mu"]), ("s", g[f"{name}/s"])):
                na, nr, nw = (x.cpu().numpy() for x in wt.upward(t(a)))
                assert np.array_equal(na, g[f"{key}/{attr}_attrs"])
                assert np.array_equal(nr, g[f"{key}/{attr}_reppoints"])
                assert np.array_equal(nw, g[f"{key}/{attr}_weights"])
            w = torch.full((p.shape[0],), float(g["width"]), dtype=torch.float64, device=DEV)
            mu, s = t(g[f"{name}/mu"]), t(g[f"{name}/s"])
            for op, a, ref in (("A_cpu", mu, g[f"{key}/A_cpu"]), ("AT", s, g[f"{key}/AT"]),
                               ("G", mu, g[f"{key}/G"])):
                out = wt.apply(op, a, w, precision="fp64").cpu().numpy()
                assert normwise(out, ref) <= TOL64, (key, op)
            out = wt.apply("A_cpu", mu, w, continuous=True, precision="fp64").cpu().numpy()
            assert normwise(out, g[f"{key}/A_cpu_cont"]) <= TOL64


def test_upward_bit_exact(cloud):
    name, p = cloud
    say("upward", name)
    rng = np.random.default_rng(5)
    wt = WindingNumberTreecode(t(p), leaf_size=16)
    ot = oracle.build_tree(p, 15, 16)
    for a in (synth.random_normals(p.shape[0], 6), rng.normal(size=(p.shape[0], 1))):
        for got, ref in zip(wt.upward(t(a)), oracle.upward(ot, p, a)):
            assert np.array_equal(got.cpu().numpy(), ref)
    wfix = rng.uniform(0.1, 2.0, p.shape[0])
    mu = synth.random_normals(p.shape[0], 7)
    for got, ref in zip(wt.upward(t(mu), t(wfix)), oracle.upward(ot, p, mu, wfix)):
        assert np.array_equal(got.cpu().numpy(), ref)


OPS = ["A", "A_cpu", "AT", "G", "A21", "A22", "A23", "A2T1", "A2T2", "A2T3", "A3"]


@pytest.mark.parametrize("leaf", [1, 32])
def test_every_operator_matches_oracle(cloud, leaf):
    name, p = cloud
    n = p.shape[0]
    rng = np.random.default_rng(leaf)
    wt = WindingNumberTreecode(t(p, torch.float32), leaf_size=leaf)
    ot = oracle.build_tree(p, 15, leaf)
    widths = rng.uniform(0.002, 0.02, n)
    for op in OPS:
        say("op", name, leaf, op)
        a = synth.random_normals(n, 8) if oracle.IN_DIM[op] == 3 else rng.normal(size=(n, 1))
        ref, (cf, cn, cv) = oracle.apply(op, ot, p, a, widths, counts=True)
        for dt in (torch.float64, torch.float32):
            out64 = wt.apply(op, t(a, dt), t(widths, dt), precision="fp64").double().cpu().numpy()
            out32 = wt.apply(op, t(a, dt), t(widths, dt), precision="fp32c").double().cpu().numpy()
            extra = 0.0 if dt == torch.float64 else 2e-7
            assert normwise(out64, ref) <= TOL64 + extra, (op, dt)
            assert normwise(out32, ref) <= TOL32C, (op, dt)
        gf, gn, gv = wt.counts(op, t(a))
        assert np.array_equal(gf.cpu().numpy(), cf)
        assert np.array_equal(gn.cpu().numpy(), cn)
        assert np.array_equal(gv.cpu().numpy(), cv)


@pytest.mark.parametrize("leaf", [1, 32])
def test_walk_pieces_match_oracle_and_whole_walks(leaf):
    """Walks cut into node-range pieces (balance with a small piece cost, so
    that most waves are cut, into up to 16 pieces each) give the oracle's sums
    for every operator in both precisions, and the uncut walks' to the
    rounding of the regrouped sums; unbalancing restores the uncut results
    bit for bit."""
    p = synth.normalize(synth.ellipsoids(40000, 5)).astype(np.float32).astype(np.float64)
    n = p.shape[0]
    rng = np.random.default_rng(30 + leaf)
    wt = WindingNumberTreecode(t(p), leaf_size=leaf)
    ot = oracle.build_tree(p, 15, leaf)
    widths = rng.uniform(0.002, 0.01, n)
    mu = synth.random_normals(n, 31)
    s3 = rng.normal(size=(n, 3))
    cases = ([(op, False) for op in OPS + ["A2", "A2T_sum"]]
             + [("A3", True), ("A", True), ("A22", True)])

    def attrs(op):
        if op in ("A2T_sum", "A2"):
            return s3 if op == "A2T_sum" else mu
        return mu if oracle.IN_DIM[op] == 3 else s3[:, :1]

    def run():
        return {(op, c, prec): wt.apply(op, t(attrs(op)), t(widths), continuous=c,
                                        precision=prec).cpu().numpy()
                for op, c in cases for prec in ("fp64", "fp32c")}

    whole = run()
    wt.balance("A2", t(mu), piece_cost=0.15, max_pieces=16)
    pieces, cut = wt.walk_pieces
    waves = (n + 63) // 64
    say("pieces", leaf, pieces, "pieces,", cut, "of", waves, "waves cut")
    assert cut > waves // 2 and pieces >= waves + cut
    cutr = run()
    for op, c in cases:
        if op in ("A2", "A2T_sum"):
            if op == "A2":
                ref = np.concatenate([oracle.apply(f"A2{k}", ot, p, mu, widths, continuous=c)
                                      for k in (1, 2, 3)], 1)
            else:
                ref = sum(oracle.apply(f"A2T{k + 1}", ot, p, s3[:, k:k + 1], widths)
                          for k in range(3))
        else:
            ref = oracle.apply(op, ot, p, attrs(op), widths, continuous=c)
        for prec, tol, same in (("fp64", TOL64, 1e-13), ("fp32c", TOL32C, 1e-6)):
            out = cutr[(op, c, prec)]
            assert normwise(out, ref) <= tol, (op, c, prec)
            assert normwise(out, whole[(op, c, prec)]) <= same, (op, c, prec)
    wt.unbalance()
    assert wt.walk_pieces == (0, 0)
    again = run()
    for k in whole:
        assert np.array_equal(again[k], whole[k]), k


@pytest.mark.parametrize("theta", [0.4, 0.6])
def test_theta_variants(theta):
    p = synth.normalize(synth.ellipsoids(4000, 9)).astype(np.float32).astype(np.float64)
    say("theta", theta)
    mu = synth.random_normals(p.shape[0], 10)
    wt = WindingNumberTreecode(t(p), leaf_size=16, theta=theta)
    ot = oracle.build_tree(p, 15, 16)
    for op in ("A21", "A3", "G"):
        ref = oracle.apply(op, ot, p, mu, 0.005, theta=theta)
        out = wt.apply(op, t(mu), t(np.full(p.shape[0], 0.005)), precision="fp64").cpu().numpy()
        assert normwise(out, ref) <= TOL64


def test_continuous_kernels(cloud):
    name, p = cloud
    mu = synth.random_normals(p.shape[0], 11)
    wt = WindingNumberTreecode(t(p), leaf_size=8)
    ot = oracle.build_tree(p, 15, 8)
    w = np.full(p.shape[0], 0.01)
    for op in ("A", "A_cpu", "A3", "A22"):
        say("continuous", name, op)
        ref = oracle.apply(op, ot, p, mu, w, continuous=True)
        for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
            out = wt.apply(op, t(mu), t(w), continuous=True, precision=prec).cpu().numpy()
            assert normwise(out, ref) <= tol, (op, prec)


def test_fused_operators_equal_reference_calls(cloud):
    name, p = cloud
    say("fused", name)
    n = p.shape[0]
    mu = synth.random_normals(n, 12)
    s3 = np.random.default_rng(13).normal(size=(n, 3))
    w = np.full(n, 0.004)
    wt = WindingNumberTreecode(t(p), leaf_size=32)
    ot = oracle.build_tree(p, 15, 32)
    refa = np.concatenate([oracle.apply(f"A2{k}", ot, p, mu, w) for k in (1, 2, 3)], 1)
    reft = sum(oracle.apply(f"A2T{k + 1}", ot, p, s3[:, k:k + 1], w) for k in range(3))
    for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
        a = wt.forward_A2(t(mu), t(w), precision=prec).cpu().numpy()
        r = wt.forward_A2T_sum(t(s3), t(w), precision=prec).cpu().numpy()
        assert normwise(a, refa) <= tol
        assert normwise(r, reft) <= tol
    # the reference-named single calls agree with the fused ones
    a21 = wt.forward_A21(t(mu), t(w)).cpu().numpy()
    assert np.array_equal(a21[:, 0], wt.forward_A2(t(mu), t(w), precision="fp64").cpu().numpy()[:, 0])


def test_external_and_grid_queries(cloud):
    name, p = cloud
    say("queries", name)
    n = p.shape[0]
    mu = synth.random_normals(n, 14) * 1e-3
    wt = WindingNumberTreecode(t(p), leaf_size=16)
    ot = oracle.build_tree(p, 15, 16)
    rng = np.random.default_rng(15)
    q = rng.uniform(-1.2, 1.2, (3000, 3))
    for op in ("A3", "A"):
        ref = oracle.apply(op, ot, p, mu, 0.01, queries=q, continuous=True)
        for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
            out = wt.query(t(q), t(mu), 0.01, op=op, continuous=True, precision=prec).cpu().numpy()
            assert normwise(out, ref) <= tol, (op, prec)
    dims, lo = (13, 9, 22), (-1.2, -1.1, -1.0)
    step = (2.4 / 12, 2.2 / 8, 2.0 / 21)
    ii = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
    gq = np.asarray(lo) + ii * np.asarray(step)
    ref = oracle.apply("A3", ot, p, mu, 0.01, queries=gq, continuous=True).reshape(dims)
    for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
        out = wt.grid_query(t(mu), lo, step, dims, 0.01, precision=prec).cpu().numpy()
        assert normwise(out, ref) <= tol


@pytest.mark.parametrize("op", ["A2", "A21", "A3", "A", "A_cpu"])
def test_exact_adjoint(op):
    p = synth.normalize(synth.ellipsoids(3000, 16)).astype(np.float32).astype(np.float64)
    n = p.shape[0]
    say("adjoint", op)
    rng = np.random.default_rng(17)
    wt = WindingNumberTreecode(t(p), leaf_size=16)
    ot = oracle.build_tree(p, 15, 16)
    wfix = rng.uniform(0.5, 1.5, n)
    widths = rng.uniform(0.003, 0.01, n)
    _, rep, _ = oracle.upward(ot, p, np.zeros((n, 3)), wfix)
    outd = 3 if op == "A2" else 1
    g = rng.normal(size=(n, outd))
    for cont in (False, True):
        if op == "A2":
            ref = sum(oracle.adjoint(f"A2{k + 1}", ot, p, g[:, k:k + 1], widths, rep, continuous=cont)
                      for k in range(3))
        else:
            ref = oracle.adjoint(op, ot, p, g, widths, rep, continuous=cont)
        for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
            out = wt.apply_adjoint(op, t(g), t(widths), t(wfix), continuous=cont,
                                   precision=prec).cpu().numpy()
            assert normwise(out, ref) <= tol, (op, cont, prec)
        # <A x, g> = <x, A^T g> with the GPU forward and the GPU adjoint (FP64)
        x = rng.normal(size=(n, 3))
        ax = wt.apply(op, t(x), t(widths), weights=t(wfix), continuous=cont,
                      precision="fp64").cpu().numpy()
        atg = wt.apply_adjoint(op, t(g), t(widths), t(wfix), continuous=cont,
                               precision="fp64").cpu().numpy()
        lhs, rhs = (ax * g).sum(), (x * atg).sum()
        assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs))


def test_degenerate_inputs():
    say("degenerate")
    for p in (np.zeros((1, 3)) + 0.25, np.repeat([[0.1, -0.2, 0.3]], 70, axis=0),
              np.random.default_rng(18).uniform(-2, 2, (500, 3)),
              np.random.default_rng(19).uniform(-1, 1, (5, 3))):
        n = p.shape[0]
        mu = synth.random_normals(n, 20)
        w = np.full(n, 0.01)
        for leaf in (1, 32):
            wt = WindingNumberTreecode(t(p), leaf_size=leaf)
            ot = oracle.build_tree(p, 15, leaf)
            for a, b in zip(wt.reference_layout(), ot.arrays()):
                assert np.array_equal(a, b)
            ref = oracle.apply("A21", ot, p, mu, w)
            out = wt.apply("A21", t(mu), t(w), precision="fp64").cpu().numpy()
            assert np.allclose(out, ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max()))
    # empty cloud: no nodes, every call is a no-op
    wt = WindingNumberTreecode(torch.zeros((0, 3), device=DEV, dtype=torch.float64))
    assert wt.info()["num_nodes"] == 0
    out = wt.apply("A2", torch.zeros((0, 3), device=DEV, dtype=torch.float64),
                   torch.zeros(0, device=DEV, dtype=torch.float64))
    assert out.shape == (0, 3)


def test_bunny_reference_result_on_gpu():
    """main_CAGR.py's schedule on the GPU operators reproduces the
    reference-held bunny normals (the file was written with An = 0)."""
    from cagr_amd.solve import PRESETS, cagr_schedule
    from oracle import solve_ref
    say("bunny")
    d = np.load(os.path.join(GOLD, "bunny.npz"))
    p = solve_ref.normalize(d["points"])
    wt = WindingNumberTreecode(t(p, torch.float64), leaf_size=1)
    normals, _ = cagr_schedule(wt, *PRESETS["l0"], an=0.0, precision="fp64")
    nrm = normals.cpu().numpy()
    ref = d["normals"]
    cos = (nrm * ref).sum(1) / np.linalg.norm(nrm, axis=1) / np.linalg.norm(ref, axis=1)
    mag = np.linalg.norm(nrm, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(cos) > 0.9999 and (cos > 0).mean() == 1.0 and abs(np.median(mag) - 1) < 0.02
    # and it equals the FP64 oracle replay of the same schedule
    # (the 40 iterations amplify rounding differences by ~1e5: FP64 vs FP64
    # with a different summation order agree to ~1e-10)
    onrm, _ = solve_ref.cagr_schedule(p, *solve_ref.PRESETS["l0"], an=0.0)
    assert normwise(nrm, onrm) <= 1e-6


def test_schedule_an1_matches_oracle_replay():
    from cagr_amd.solve import cagr_schedule
    from oracle import solve_ref
    say("schedule an=1")
    p = synth.normalize(synth.sphere(2000, 21)).astype(np.float32).astype(np.float64)
    wt = WindingNumberTreecode(t(p), leaf_size=1)
    nrm, res = cagr_schedule(wt, 0.01, 0.04, precision="fp64")
    onrm, ores = solve_ref.cagr_schedule(p, 0.01, 0.04)
    assert normwise(nrm.cpu().numpy(), onrm) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_orient_end_to_end_reproduces_bunny(tmp_path, precision):
    """cagr_amd.pcio.orient -- read .xyz, normalize, float32 points on the GPU
    (main_CAGR.py:47), the 40-iteration schedule and 8 G steps, write .xyz --
    reproduces the reference-held bunny normals (written with An = 0)."""
    from cagr_amd import pcio
    say("orient bunny", precision)
    d = np.load(os.path.join(GOLD, "bunny.npz"))
    src = str(tmp_path / "bunny.xyz")
    np.savetxt(src, d["points"])
    out_dir = str(tmp_path / "out")
    nrm, pn = pcio.orient(src, "l0", an=0.0, precision=precision, out_dir=out_dir)
    assert np.array_equal(pn, pcio.normalize(d["points"])[0])
    ref = d["normals"]
    cos = (nrm * ref).sum(1) / np.linalg.norm(nrm, axis=1) / np.linalg.norm(ref, axis=1)
    mag = np.linalg.norm(nrm, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(cos) > 0.9999 and (cos > 0).mean() == 1.0 and abs(np.median(mag) - 1) < 0.02
    pts, nrm2 = pcio.load_points(os.path.join(out_dir, "bunny.xyz"))
    assert np.array_equal(pts, d["points"]) and np.array_equal(nrm2, nrm.astype(np.float64))


@pytest.mark.parametrize("leaf", [1, 16])
def test_fp64_points_on_the_cutoff_match_oracle(leaf):
    """FP64 points that are not FP32 values, in close pairs 5e-4 to 1e-3
    apart, each point's width its partner's FP64 distance times 1 + d with
    |d| <= 2e-5.  The float copies of the points move a pair's distance by up
    to ~1e-4 of itself, so FP32 arithmetic alone would decide many of these
    cutoff tests wrongly, and a wrong decision adds or drops the row's
    largest term (the kernels are zero inside the width).  The FP32c results
    match the oracle only if every such test is redone in FP64 on the FP64
    points: the near pairs (leaf 16) and the single-point far nodes (leaf 1),
    in the forward operators, the fused forms and both adjoints.

    Bar 1e-3 normwise: a wrong decision moves a row by about its largest term
    (errors ~0.01-1 of the largest entry), while the FP32c form computes d
    from the float copies, which moves the dominant terms by up to
    3 * 2 * 2^-25 / 5e-4 ~ 4e-4 of themselves (measured 1.0e-4)."""
    say("fp64 cutoff", leaf)
    rng = np.random.default_rng(31)
    m = 1500
    base = synth.normalize(synth.ellipsoids(m, 12))
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    p = np.concatenate([base, base + u * rng.uniform(5e-4, 1e-3, m)[:, None]])
    assert not np.array_equal(p, p.astype(np.float32).astype(np.float64))
    n = 2 * m
    d = p - p[np.concatenate([np.arange(m, n), np.arange(m)])]
    widths = np.sqrt((d[:, 0] ** 2 + d[:, 1] ** 2) + d[:, 2] ** 2) * (
        1.0 + rng.uniform(-2e-5, 2e-5, n))
    wt = WindingNumberTreecode(t(p), leaf_size=leaf)
    ot = oracle.build_tree(p, 15, leaf)
    mu = synth.random_normals(n, 5)
    s3 = rng.normal(size=(n, 3))
    for op in ("A", "A21", "A2T1", "AT"):
        a = mu if oracle.IN_DIM[op] == 3 else s3[:, :1]
        ref = oracle.apply(op, ot, p, a, widths)
        out = wt.apply(op, t(a), t(widths), precision="fp32c").cpu().numpy()
        assert normwise(out, ref) <= TOLF64PTS, (op, normwise(out, ref))
    ref_a = np.concatenate([oracle.apply(f"A2{k}", ot, p, mu, widths) for k in (1, 2, 3)], 1)
    ref_t = sum(oracle.apply(f"A2T{k + 1}", ot, p, s3[:, k:k + 1], widths) for k in range(3))
    a = wt.forward_A2(t(mu), t(widths), precision="fp32c").cpu().numpy()
    r = wt.forward_A2T_sum(t(s3), t(widths), precision="fp32c").cpu().numpy()
    assert normwise(a, ref_a) <= TOLF64PTS, normwise(a, ref_a)
    assert normwise(r, ref_t) <= TOLF64PTS, normwise(r, ref_t)
    wfix = rng.uniform(0.5, 1.5, n)
    _, rep, _ = oracle.upward(ot, p, np.zeros((n, 3)), wfix)
    g = rng.normal(size=(n, 1))
    ref = oracle.adjoint("A", ot, p, g, widths, rep)
    out = wt.apply_adjoint("A", t(g), t(widths), t(wfix), precision="fp32c").cpu().numpy()
    assert normwise(out, ref) <= TOLF64PTS, normwise(out, ref)
    wt.adjoint_plan(t(wfix))
    out = wt.apply_adjoint_planned("A", t(g), t(widths), precision="fp32c").cpu().numpy()
    assert normwise(out, ref) <= TOLF64PTS, normwise(out, ref)


@pytest.mark.parametrize("leaf", [16, 32, 64])
def test_dense_cloud_small_width_matches_oracle(leaf):
    """A cloud as dense as the benchmark's (80k points on ellipsoids) with its
    width 0.002: most near leaves then have few near lanes and need no cutoff
    test, the case where the walks spread a leaf's points over idle lanes.
    Rows of the fused A2 and A2T_sum at 3000 sampled points against the oracle
    evaluated at those points (the oracle's queries are plain points, so this
    is the same sum)."""
    say("dense", leaf)
    p = synth.normalize(synth.ellipsoids(80000, 7)).astype(np.float32).astype(np.float64)
    n = p.shape[0]
    rng = np.random.default_rng(leaf)
    idx = np.sort(rng.choice(n, 3000, replace=False))
    mu = synth.random_normals(n, 3)
    s3 = rng.normal(size=(n, 3))
    w = np.full(n, 0.002)
    wt = WindingNumberTreecode(t(p, torch.float32), leaf_size=leaf)
    ot = oracle.build_tree(p, 15, leaf)
    ref_a = np.concatenate([oracle.apply(f"A2{k}", ot, p, mu, w[idx], queries=p[idx])
                            for k in (1, 2, 3)], 1)
    ref_t = sum(oracle.apply(f"A2T{k + 1}", ot, p, s3[:, k:k + 1], w[idx], queries=p[idx])
                for k in range(3))
    for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
        a = wt.forward_A2(t(mu), t(w), precision=prec).cpu().numpy()[idx]
        r = wt.forward_A2T_sum(t(s3), t(w), precision=prec).cpu().numpy()[idx]
        assert normwise(a, ref_a) <= tol, (prec, normwise(a, ref_a))
        assert normwise(r, ref_t) <= tol, (prec, normwise(r, ref_t))


@pytest.mark.parametrize("an", [0.5, 1.0, 2.0])
def test_continuous_branch_carries_weight(an):
    """The continuous inner kernel of A2k / A3 (inside the width:
    e^{An w} [d.mu/w^3 + An/2 mu_k/w], cuda_kernels.cu:140-147, 179-185,
    216-222) at several convection strengths, with a width so large that
    every query has tens of sources inside it: the inner branch then carries
    a large share of every output (asserted), so a wrong exponent or a dropped
    An term there cannot hide behind the far field.  Self queries, explicit
    queries (GaussRecon's sample pass) and a grid, both precisions."""
    p = synth.normalize(synth.sphere(2500, 21)).astype(np.float32).astype(np.float64)
    n = p.shape[0]
    mu = synth.random_normals(n, 22)
    w = np.full(n, 0.3)
    wt = WindingNumberTreecode(t(p), leaf_size=8)
    ot = oracle.build_tree(p, 15, 8)
    for op in ("A21", "A22", "A23", "A3"):
        say("continuous branch", an, op)
        ref = oracle.apply(op, ot, p, mu, w, continuous=True, an=an)
        outer = oracle.apply(op, ot, p, mu, w, continuous=False, an=an)
        assert normwise(ref, outer) > 0.2, "the inner branch must carry weight"
        for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
            out = wt.apply(op, t(mu), t(w), continuous=True, an=an, precision=prec).cpu().numpy()
            assert normwise(out, ref) <= tol, (op, prec, normwise(out, ref))
    fused = oracle.apply("A21", ot, p, mu, w, continuous=True, an=an)
    out = wt.apply("A2", t(mu), t(w), continuous=True, an=an, precision="fp64").cpu().numpy()
    assert normwise(out[:, :1], fused) <= TOL64
    q = np.random.default_rng(23).uniform(-0.6, 0.6, (2000, 3)) + p[:2000]
    ref = oracle.apply("A3", ot, p, mu, 0.3, queries=q, continuous=True, an=an)
    for prec, tol in (("fp64", TOL64), ("fp32c", TOL32C)):
        out = wt.query(t(q), t(mu), 0.3, op="A3", precision=prec, an=an).cpu().numpy()
        assert normwise(out, ref) <= tol, (prec, normwise(out, ref))
    dims, lo, step = (9, 11, 10), (-1.0, -1.0, -1.0), (0.25, 0.2, 0.22)
    ii = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
    gq = np.asarray(lo) + ii * np.asarray(step)
    ref = oracle.apply("A3", ot, p, mu, 0.3, queries=gq, continuous=True, an=an).reshape(dims)
    out = wt.grid_query(t(mu), lo, step, dims, 0.3, precision="fp64", an=an).cpu().numpy()
    assert normwise(out, ref) <= TOL64


def test_calls_on_two_streams_are_ordered():
    """One handle used from two streams with no synchronisation between the
    calls: every call waits for the previous one's work on the handle's shared
    records, so the results equal the same calls made on one stream."""
    p = synth.normalize(synth.ellipsoids(60000, 24)).astype(np.float32)
    n = p.shape[0]
    say("two streams", n)
    wt = WindingNumberTreecode(t(p, torch.float32), leaf_size=32)
    mus = [t(synth.random_normals(n, 25 + i), torch.float32) for i in range(4)]
    w = t(np.full(n, 0.002), torch.float32)
    want = [wt.forward_A2(m, w, precision="fp32c").clone() for m in mus]
    want_t = [wt.forward_A2T_sum(m, w, precision="fp32c").clone() for m in mus]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    got, got_t = [], []
    for i, m in enumerate(mus):
        with torch.cuda.stream(s1 if i % 2 == 0 else s2):
            got.append(wt.forward_A2(m, w, precision="fp32c"))
        with torch.cuda.stream(s2 if i % 2 == 0 else s1):
            got_t.append(wt.forward_A2T_sum(m, w, precision="fp32c"))
    torch.cuda.synchronize()
    for a, b in zip(got + got_t, want + want_t):
        assert torch.equal(a, b)
