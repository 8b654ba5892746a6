"""GPU tests of the batched form (one forest, one persistent launch for all
clouds) and of its autotuner.  Every cloud of a forest must give exactly what
its own single-cloud tree gives: the same topology (node and point ids offset
by the cloud's position), the same interaction counts, the same FP64 results
(to rounding) as the oracle run on that cloud alone -- whatever the launch
shape and the order of the work queue."""
import numpy as np
import pytest
import torch

import oracle
from cagr_amd import WindingNumberTreecode, synth
from cagr_amd.autotune import autotune
from cagr_amd.batched import BatchedTreecode

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def say(*a):
    print("[case]", *a, flush=True)


def make_clouds(seed=0, count=7):
    rng = np.random.default_rng(seed)
    sizes = [1, 0, 3000, 700, 64, 4100, 33][:count]
    clouds = []
    for i, n in enumerate(sizes):
        if n == 0:
            clouds.append(np.zeros((0, 3)))
            continue
        kind = i % 3
        p = [synth.sphere, synth.torus, synth.ellipsoids][kind](n, seed + i)
        clouds.append(synth.normalize(p) if n > 1 else rng.uniform(-0.5, 0.5, (1, 3)))
    return [c.astype(np.float32).astype(np.float64) for c in clouds]


@pytest.mark.parametrize("leaf", [1, 32])
def test_forest_topology_equals_single_trees(leaf):
    say("forest topology", leaf)
    clouds = make_clouds()
    bt = BatchedTreecode([torch.from_numpy(c).to(DEV) for c in clouds], leaf_size=leaf)
    lay = bt.reference_layout()
    noff = bt.node_offsets()
    poff = bt.offsets
    for c, p in enumerate(clouds):
        ot = oracle.build_tree(p, 15, leaf)
        r0, r1 = noff[c], noff[c + 1]
        assert r1 - r0 == ot.num_nodes
        if ot.num_nodes == 0:
            continue
        par = lay[0][r0:r1].copy()
        par[par >= 0] -= r0
        assert np.array_equal(par, ot.parent)
        ch = lay[1][r0:r1].copy()
        ch[ch >= 0] -= r0
        assert np.array_equal(ch, ot.children)
        assert np.array_equal(lay[2][r0:r1], ot.is_leaf)
        assert np.array_equal(lay[3][r0:r1], ot.half_w)
        assert np.array_equal(lay[4][r0:r1], ot.num_points)
        s0 = lay[6][r0]
        ll = ot.node2point_index.shape[0]
        assert np.array_equal(lay[6][r0:r1] - s0, ot.node2point_indexstart)
        assert np.array_equal(lay[5][s0:s0 + ll] - poff[c], ot.node2point_index)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-6)])
def test_forest_operators_equal_per_cloud_oracle(prec, tol):
    say("forest ops", prec)
    clouds = make_clouds(1)
    leaf = 16
    pts = [torch.from_numpy(c).to(DEV) for c in clouds]
    bt = BatchedTreecode(pts, leaf_size=leaf)
    n = bt.points.shape[0]
    rng = np.random.default_rng(2)
    mu = synth.random_normals(n, 3)
    s3 = rng.normal(size=(n, 3))
    w = rng.uniform(0.003, 0.02, n)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    outs = {}
    for wpb in (1, 4, 8):
        bt.set_launch(wpb)
        for balanced in (False, True):
            if balanced:
                bt.balance("A2", T(mu))
            else:
                bt.unbalance()
            a = bt.forward_A2(T(mu), T(w), precision=prec).cpu().numpy()
            r = bt.forward_A2T_sum(T(s3), T(w), precision=prec).cpu().numpy()
            g = bt.forward_G(T(mu), T(w)).cpu().numpy() if prec == "fp64" else None
            outs[(wpb, balanced)] = (a, r, g)
    base = outs[(1, False)]
    for k, v in outs.items():
        if prec == "fp64":      # same decisions and same per-query order: identical
            assert np.array_equal(v[0], base[0]) and np.array_equal(v[1], base[1]), k
    a, r, g = base
    for c, p in enumerate(clouds):
        lo, hi = bt.offsets[c], bt.offsets[c + 1]
        if hi == lo:
            continue
        ot = oracle.build_tree(p, 15, leaf)
        refa = np.concatenate([oracle.apply(f"A2{k}", ot, p, mu[lo:hi], w[lo:hi]) for k in (1, 2, 3)], 1)
        reft = sum(oracle.apply(f"A2T{k + 1}", ot, p, s3[lo:hi, k:k + 1], w[lo:hi]) for k in range(3))
        sa = max(np.abs(refa).max(), 1e-300)
        st = max(np.abs(reft).max(), 1e-300)
        assert np.abs(a[lo:hi] - refa).max() <= tol * sa, c
        assert np.abs(r[lo:hi] - reft).max() <= tol * st, c
        if g is not None:
            refg = oracle.apply("G", ot, p, mu[lo:hi], w[lo:hi])
            assert np.abs(g[lo:hi] - refg).max() <= tol * max(np.abs(refg).max(), 1e-300)
        _, (cf, cn, cv) = oracle.apply("A21", ot, p, mu[lo:hi], w[lo:hi], counts=True)
        gf, gn, gv = (x.cpu().numpy()[lo:hi] for x in bt.counts("A2", T(mu)))
        assert np.array_equal(gf, cf) and np.array_equal(gn, cn) and np.array_equal(gv, cv)


def test_persistent_single_tree_equals_regular_launch():
    say("persistent single")
    p = synth.normalize(synth.ellipsoids(20000, 4)).astype(np.float32)
    pt = torch.from_numpy(p).to(DEV)
    wt = WindingNumberTreecode(pt, leaf_size=32, precision="fp64")
    mu = torch.from_numpy(synth.random_normals(p.shape[0], 5).astype(np.float32)).to(DEV)
    w = torch.full((p.shape[0],), 0.004, device=DEV)
    a0 = wt.forward_A2(mu, w)
    r0 = wt.forward_A2T_sum(a0, w)
    wt.L.cagr_set_launch(wt._h, 1, 2, 0)
    a1 = wt.forward_A2(mu, w)
    r1 = wt.forward_A2T_sum(a0, w)
    assert torch.equal(a0, a1) and torch.equal(r0, r1)


def test_autotune_small_batch():
    say("autotune")
    clouds = [synth.normalize(p) for p in (synth.sphere(3000, 1), synth.torus(2500, 2),
                                              synth.ellipsoids(4000, 3))]
    pts = [torch.from_numpy(c.astype(np.float32)).to(DEV) for c in clouds]
    mus = [torch.from_numpy(synth.random_normals(c.shape[0], 9).astype(np.float32)).to(DEV)
           for c in clouds]
    rep = autotune(pts, mus, 0.004, waves_per_block=(2, 4), steps=2, rounds=2,
                   samples_per_cloud=32, log=print)
    assert len(rep.errors) == 9
    admissible = {lt for lt, e in rep.errors.items() if e <= rep.budget}
    assert {(r.leaf_size, r.theta) for r in rep.table} == admissible
    assert len(rep.table) == 2 * len(admissible)
    assert rep.best.err <= rep.budget
    assert rep.best.ms_per_pair == min(r.ms_per_pair for r in rep.table)
    # a budget nothing meets: the most accurate configuration
    tight = autotune(pts, mus, 0.004, waves_per_block=(4,), steps=1, rounds=1,
                     samples_per_cloud=32, budget=1e-12)
    assert (tight.best.leaf_size, tight.best.theta) == min(tight.errors, key=tight.errors.get)


def test_autotune_64_clouds_and_size_classes():
    """The tuner on the 64-cloud workload: its choice is within the stated
    budget against direct sums, and the tuned forest -- and the per-size-class
    forests -- give each cloud what the oracle gives on that cloud's own tree
    at the tuned leaf size and theta."""
    from cagr_amd.autotune import ClassBatches, autotune_classes
    clouds_np = [c.astype(np.float32).astype(np.float64) for c in synth.batch_clouds(64)]
    pts = [torch.from_numpy(c).to(DEV, torch.float32) for c in clouds_np]
    mus_np = [synth.random_normals(c.shape[0], 20 + i) for i, c in enumerate(clouds_np)]
    mus = [torch.from_numpy(m).to(DEV, torch.float32) for m in mus_np]
    say("autotune 64 clouds")
    rep = autotune(pts, mus, 0.002, waves_per_block=(2, 8), steps=2, rounds=2,
                   samples_per_cloud=16, log=print)
    b = rep.best
    print(f"[tuned] leaf {b.leaf_size} theta {b.theta} wpb {b.waves_per_block} "
          f"{b.ms_per_pair:.3f} ms err {b.err:.2e} stable {rep.stable}", flush=True)
    assert b.err <= rep.budget
    classes = autotune_classes(pts, mus, 0.002, waves_per_block=(4,), steps=1, rounds=1,
                               samples_per_cloud=16)
    assert sorted(np.concatenate([sel for sel, _ in classes]).tolist()) == list(range(64))
    cb = ClassBatches(pts, classes)
    mu = torch.cat(mus)
    w = torch.full((mu.shape[0],), 0.002, device=DEV)
    out = cb.forward_A2(mu, w, precision="fp64").double().cpu().numpy()
    offs = np.concatenate([[0], np.cumsum([c.shape[0] for c in clouds_np])])
    for sel, r in classes:
        for c in sel[:: max(1, len(sel) // 4)]:       # a few clouds of every class
            p = clouds_np[c]
            ot = oracle.build_tree(p, 15, r.best.leaf_size)
            idx = np.sort(np.random.default_rng(c).choice(p.shape[0], 32, replace=False))
            mu32 = mus_np[c].astype(np.float32).astype(np.float64)
            ref = np.concatenate([oracle.apply(f"A2{k}", ot, p, mu32, 0.002, queries=p[idx],
                                               theta=r.best.theta) for k in (1, 2, 3)], 1)
            got = out[offs[c]:offs[c + 1]][idx]
            assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max(), c
