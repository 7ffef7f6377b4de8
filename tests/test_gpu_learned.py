"""Learned Widom proposal on the GPU: network forward and gradient against the
FP64 oracle (with the kernels' BF16 rounding points), descriptors against
brute force, unbiasedness of the importance-weighted estimate, and the gain of
the trained network over uniform insertion."""
import math

import numpy as np
import pytest

from hsmc_amd import gpu
from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ctx_with(n, rho, sweeps, seed=1):
    o = O.System(2, n, rho=rho, neigh_dr=1.0, max_part=12, seed=seed)
    L = o.box()["lx"]
    conf = o.config()
    ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.12, seed=seed)])
    ctx.upload(conf)
    ctx.sweeps(sweeps)
    return ctx, L


def test_mlp_forward_matches_oracle():
    ctx, _ = _ctx_with(4, 0.5, 0)
    rng = np.random.default_rng(1)
    p = (rng.standard_normal(M.NPARAM) * 0.2).astype(np.float32)
    net = gpu.Mlp(ctx)
    net.set(p)
    assert np.array_equal(net.get(), p)
    x = rng.standard_normal((1000, 12)).astype(np.float32)
    y = net.forward(x)
    ref, _ = M.forward(p.astype(np.float64), x.astype(np.float64), bf=True)
    assert np.max(np.abs(y - ref)) <= 2e-3 * max(1.0, np.max(np.abs(ref)))


def test_mlp_gradient_matches_oracle():
    ctx, _ = _ctx_with(4, 0.5, 0)
    rng = np.random.default_rng(2)
    p = (rng.standard_normal(M.NPARAM) * 0.2).astype(np.float32)
    net = gpu.Mlp(ctx)
    net.set(p)
    x = rng.standard_normal((777, 12)).astype(np.float32)   # not a multiple of 64
    t = rng.standard_normal(777).astype(np.float32)
    loss, g = net.gradient(x, t)
    rl, rg = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss - rl) <= 1e-3 * rl
    scale = np.max(np.abs(rg))
    for name, a, b in (("W1", M.OW1, M.OB1), ("b1", M.OB1, M.OW2), ("W2", M.OW2, M.OB2),
                       ("b2", M.OB2, M.OW3), ("w3", M.OW3, M.OB3), ("b3", M.OB3, M.NPARAM)):
        err = np.max(np.abs(g[a:b] - rg[a:b]))
        assert err <= 5e-3 * scale, (name, err, scale)
    # and against the unrounded FP64 gradient, to the BF16 resolution
    _, g64 = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert np.linalg.norm(g - g64) <= 0.05 * np.linalg.norm(g64)


def test_descriptors_match_brute_force():
    ctx, L = _ctx_with(6, 0.8, 20)
    conf = ctx.download()
    rng = np.random.default_rng(3)
    pts = rng.random((300, 3)) * L
    d = gpu.point_descriptors(ctx, pts)
    ref = M.nearest12(conf, L, pts)
    # exact up to FP32 rounding whenever the 12 neighbours lie within the
    # 1.5-sigma descriptor grid (always at this density)
    assert np.max(np.abs(d - ref)) < 1e-5


def _uniform_p(ctx, n_per, samples, seed):
    h = [ctx.widom_uniform(seed, s, 0, n_per) for s in range(samples)]
    f = np.array(h) / n_per
    return f.mean(), f.std(ddof=1) / math.sqrt(samples)


def test_learned_estimator_is_unbiased():
    # rho = 0.5: P_ins ~ 2%, checkable against plain uniform insertion;
    # an untrained (random) network and a trained one must both agree
    ctx, L = _ctx_with(8, 0.5, 50)
    pu, su = _uniform_p(ctx, 200000, 20, 11)
    for trained in (False, True):
        net = gpu.Mlp(ctx, seed=5)
        if trained:
            net.train(0.25, 99, 0, 200000, epochs=3, batch=4096, lr=2e-3, beta=10.0)
        est = [net.widom(0.25, 0.9, 21, s, 200000)["estimate"] for s in range(20)]
        pl, sl = np.mean(est), np.std(est, ddof=1) / math.sqrt(len(est))
        assert abs(pl - pu) < 4 * math.hypot(su, sl), (trained, pl, pu, su, sl)
        assert abs(pl / pu - 1) < 0.05


def test_trained_network_beats_uniform():
    # rho = 0.8, N = 4000: a trained proposal lowers the variance per insertion
    ctx, L = _ctx_with(10, 0.8, 300)
    n = 400000
    pu, su = _uniform_p(ctx, n, 16, 31)
    net = gpu.Mlp(ctx, seed=9)
    loss = net.train(0.25, 77, 0, 400000, epochs=6, batch=4096, lr=2e-3, beta=20.0)
    assert math.isfinite(loss)
    est = [net.widom(0.25, 0.9, 41, s, n)["estimate"] for s in range(16)]
    pl, sl = np.mean(est), np.std(est, ddof=1) / math.sqrt(len(est))
    assert abs(pl - pu) < 4 * math.hypot(su, sl), (pl, pu, su, sl)
    gain = (su / sl) ** 2   # variance ratio at equal insertion counts
    print(f"rho=0.8 P_ins uniform {pu:.4e}+-{su:.1e} learned {pl:.4e}+-{sl:.1e} gain {gain:.1f}")
    assert gain > 2.0


def test_systematic_sampler_weights():
    """mix = 0 makes q uniform (q_v = 1/V): every insertion's weight is then
    exactly 1/n, so the estimate equals hits / n, and all n insertions are
    made whatever n is."""
    ctx, L = _ctx_with(6, 0.5, 20)
    net = gpu.Mlp(ctx, seed=3)
    for n in (1, 777, 100000):
        r = net.widom(0.25, 0.0, 5, 1, n)
        assert r["n"] == n
        assert abs(r["estimate"] - r["hits"] / n) <= 1e-12 * max(1.0, r["hits"] / n)
    # with a proposal that is not uniform the count is still exactly n
    r = net.widom(0.25, 0.9, 5, 2, 54321)
    assert r["n"] == 54321 and r["hits"] > 0


@pytest.mark.parametrize("n,rho", [(10, 0.8), (3, 0.6)])
def test_voxel_scores_match_point_path(n, rho):
    """Voxel scores (tiled kernel for the N = 4000 box of edge 17.1, per-voxel
    kernel for the N = 108 box of edge 5.65, below the 6.5 tiles need) equal the network
    applied to descriptors taken at the voxel centres by the per-point FP64
    path, up to the rounding of FP32 distances feeding BF16 inputs."""
    ctx, L = _ctx_with(n, rho, 100)
    net = gpu.Mlp(ctx, seed=4)
    net.train(0.25, 5, 0, 100000, epochs=2, batch=4096, lr=2e-3, beta=10.0)
    scores = net.voxel_scores(0.25)
    m = int(math.ceil(L / 0.25))
    assert len(scores) == m ** 3
    v = np.random.default_rng(2).choice(m ** 3, size=min(5000, m ** 3), replace=False)
    side = L / m
    ix, iy, iz = v // (m * m), (v // m) % m, v % m
    centres = np.stack([(ix + 0.5) * side, (iy + 0.5) * side, (iz + 0.5) * side], axis=1)
    ref = net.forward(gpu.point_descriptors(ctx, centres))
    diff = np.abs(scores[v] - ref)
    scale = max(1.0, float(np.max(np.abs(ref))))
    assert np.quantile(diff, 0.99) <= 1e-4 * scale, np.quantile(diff, 0.99)
    assert diff.max() <= 0.05 * scale, diff.max()
    ctx.close()


def _min_image(d, L):
    # compute_dist's minimum image (moves.c:419-424), per axis
    return np.where(d > L / 2.0, d - L, np.where(d < -L / 2.0, d + L, d))


def _anisotropic_ctx():
    """fcc 8 x 6 x 7 cells at rho = 0.8 (three different box edges), melted
    by 300 sweeps on the GPU."""
    lat, box = gpu.lattice(2, (8, 6, 7), 0.8)
    ctx = gpu.Context([dict(n=len(lat), box=box, dr_max=0.12, seed=7)])
    ctx.upload(lat)
    ctx.sweeps(300)
    return ctx, tuple(float(b) for b in box)


@pytest.mark.parametrize("case", ["cube-0.8", "cube-0.6", "anisotropic-0.8"])
def test_proposal_excludes_exactly_the_blocked_voxels(case):
    """The learned proposal scores only voxels that can hold a free point.
    The voxels it marks -inf are exactly those whose centre lies closer than
    1 - (half the voxel diagonal) to a sphere centre, with the squared
    distance in FP64 as compute_dist evaluates it (checked against every
    sphere within 1 of each centre, found by a periodic k-d tree); the count
    of free voxels matches; every free voxel's score is the network applied
    to the per-point descriptor at its centre; and random points inside
    excluded voxels all overlap a sphere (the exclusion radius is
    geometrically safe).  Cubic boxes at rho = 0.8 and 0.6, and a box with
    three different edges (so three voxel counts and edges)."""
    from scipy.spatial import cKDTree
    if case == "anisotropic-0.8":
        ctx, box = _anisotropic_ctx()
    else:
        n, rho = {"cube-0.8": (8, 0.8), "cube-0.6": (6, 0.6)}[case]
        ctx, L = _ctx_with(n, rho, 200)
        box = (L, L, L)
    net = gpu.Mlp(ctx, seed=4)
    net.train(0.25, 5, 0, 100000, epochs=2, batch=4096, lr=2e-3, beta=10.0)
    net.voxel_scores(0.25)   # leaves finite scores in the shared buffer first
    f, n_free = net.proposal_scores(0.25)
    m = [int(math.ceil(b / 0.25)) for b in box]
    side = [b / k for b, k in zip(box, m)]
    assert len(f) == m[0] * m[1] * m[2]
    h = 0.5 * math.sqrt(side[0] * side[0] + side[1] * side[1] + side[2] * side[2])
    r2 = (1.0 - h - 1e-9) ** 2
    conf = ctx.download()[:, 1:4]
    axes = [(np.arange(k) + 0.5) * sd for k, sd in zip(m, side)]
    centres = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    Lv = np.array(box)
    tree = cKDTree(np.mod(conf, Lv), boxsize=Lv)
    near = tree.query_ball_point(np.mod(centres, Lv), r=1.0)
    blocked = np.zeros(len(centres), dtype=bool)
    for v, js in enumerate(near):
        if js:
            d = np.stack([_min_image(centres[v][k] - conf[js, k], box[k]) for k in range(3)], axis=1)
            d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
            blocked[v] = bool((d2 < r2).any())
    assert np.array_equal(np.isneginf(f), blocked)
    assert n_free == int((~blocked).sum()) and 0 < n_free < len(f)
    free = np.flatnonzero(~blocked)
    ref = net.forward(gpu.point_descriptors(ctx, centres[free]))
    diff = np.abs(f[free] - ref)
    scale = max(1.0, float(np.max(np.abs(ref))))
    assert np.quantile(diff, 0.99) <= 1e-4 * scale and diff.max() <= 0.05 * scale
    rng = np.random.default_rng(3)
    ex = rng.choice(np.flatnonzero(blocked), size=2000, replace=False)
    pts = centres[ex] + (rng.random((2000, 3)) - 0.5) * np.array(side)
    dmin, _ = tree.query(np.mod(pts, Lv), k=1)
    assert (dmin < 1.0).all()
    ctx.close()


def test_proposal_scores_every_voxel_when_voxels_are_large():
    """A voxel of edge 1.3 has a half diagonal above 1: no voxel can be
    excluded, and every voxel gets the network's score."""
    ctx, L = _ctx_with(6, 0.6, 50)
    net = gpu.Mlp(ctx, seed=4)
    f, n_free = net.proposal_scores(1.3)
    assert n_free == len(f) and np.isfinite(f).all()
    raw = net.voxel_scores(1.3)
    assert np.array_equal(f, raw)
    ctx.close()


def test_proposal_with_no_free_voxel():
    """fcc at rho = 1.2 (nearest neighbours 1.056 apart): every voxel centre
    lies within 0.65 of a sphere, below the exclusion radius 0.784 of voxels
    of edge 0.249, so no voxel is free; every point of the box lies within
    0.75 of a sphere, so no insertion can succeed.  The proposal keeps the
    uniform part alone: all n insertions are made, none hits, and the
    estimate is exactly 0 at every mixing weight.  A recomputation after
    forget() gives the cached call's result."""
    lat, box = gpu.lattice(2, (4, 4, 4), 1.2)
    ctx = gpu.Context([dict(n=len(lat), box=box, dr_max=0.05, seed=3)])
    ctx.upload(lat)
    net = gpu.Mlp(ctx, seed=2)
    f, n_free = net.proposal_scores(0.25)
    assert len(f) == 24 ** 3 and n_free == 0 and np.isneginf(f).all()
    for mix in (0.0, 0.5, 0.9):
        r = net.widom(0.25, mix, 7, 1, 100000)
        assert r["n"] == 100000 and r["hits"] == 0 and r["estimate"] == 0.0
    assert ctx.widom_uniform(7, 1, 0, 100000) == 0
    # a free configuration after sweeps at a lower density: the cached
    # estimate equals a cold recomputation bit for bit
    ctx2, _ = _ctx_with(6, 0.6, 30)
    net2 = gpu.Mlp(ctx2, seed=2)
    first = net2.widom(0.25, 0.9, 7, 1, 100000)
    cached = net2.widom(0.25, 0.9, 7, 1, 100000)
    net2.forget()
    cold = net2.widom(0.25, 0.9, 7, 1, 100000)
    assert first == cached == cold and first["hits"] > 0
    ctx2.close()
    ctx.close()


def test_two_networks_share_a_context():
    """Two networks on one context alternate learned Widom calls with
    voxel_scores calls, all of which write the context's shared score and
    cumulative-mass buffers: each network's estimate must equal the one it
    gave from a cold cache (same seed and sample, bit for bit), never one
    computed from the other network's scores."""
    ctx, L = _ctx_with(8, 0.6, 30)
    a, b = gpu.Mlp(ctx, seed=5), gpu.Mlp(ctx, seed=6)
    b.train(0.25, 3, 0, 50000, epochs=1, batch=4096, lr=2e-3, beta=10.0)
    cold_a = a.widom(0.25, 0.9, 21, 1, 100000)["estimate"]
    cold_b = b.widom(0.25, 0.9, 21, 1, 100000)["estimate"]
    assert cold_a != cold_b   # different networks, different proposals
    for _ in range(2):
        assert a.widom(0.25, 0.9, 21, 1, 100000)["estimate"] == cold_a
        assert b.widom(0.25, 0.9, 21, 1, 100000)["estimate"] == cold_b
        b.voxel_scores(0.25)
        assert a.widom(0.25, 0.9, 21, 1, 100000)["estimate"] == cold_a
        a.voxel_scores(0.25)
        assert b.widom(0.25, 0.9, 21, 1, 100000)["estimate"] == cold_b
    # a different mixing weight rebuilds the cumulative mass, and returning
    # to the first weight still gives the cold estimate
    a.widom(0.25, 0.5, 21, 1, 100000)
    assert b.widom(0.25, 0.9, 21, 1, 100000)["estimate"] == cold_b
    assert a.widom(0.25, 0.9, 21, 1, 100000)["estimate"] == cold_a


def test_learned_widom_at_full_size():
    """The learned proposal at its own size: N = 108000 (fcc 4*30^3) at
    rho = 0.80, 206^3 voxels.  On one configuration the learned estimate of
    the insertion probability agrees with uniform insertion within 3
    combined standard errors at the same number of insertions, and the tiled
    voxel scores equal the network applied to per-point FP64 descriptors at
    the voxel centres on a sample of voxels."""
    ctx, L = _ctx_with(30, 0.80, 400, seed=3)
    n_per, samples = 20_000_000, 16
    pu, su = _uniform_p(ctx, n_per, samples, 77)
    net = gpu.Mlp(ctx, seed=12)
    net.train(0.25, 55, 0, 1_000_000, epochs=4, batch=4096, lr=2e-3, beta=20.0)
    est = [net.widom(0.25, 0.9, 91, s, n_per)["estimate"] for s in range(samples)]
    pl, sl = float(np.mean(est)), float(np.std(est, ddof=1) / math.sqrt(samples))
    print(f"N=108000 rho=0.8: P_ins uniform {pu:.5e}+-{su:.1e} learned {pl:.5e}+-{sl:.1e}")
    assert pu > 0 and sl > 0
    assert abs(pl - pu) < 3 * math.hypot(su, sl), (pl, pu, su, sl)
    scores = net.voxel_scores(0.25)
    m = int(math.ceil(L / 0.25))
    assert m == 206 and len(scores) == m ** 3
    v = np.random.default_rng(8).choice(m ** 3, size=20000, replace=False)
    side = L / m
    ix, iy, iz = v // (m * m), (v // m) % m, v % m
    centres = np.stack([(ix + 0.5) * side, (iy + 0.5) * side, (iz + 0.5) * side], axis=1)
    ref = net.forward(gpu.point_descriptors(ctx, centres))
    diff = np.abs(scores[v] - ref)
    scale = max(1.0, float(np.max(np.abs(ref))))
    assert np.quantile(diff, 0.99) <= 1e-4 * scale, np.quantile(diff, 0.99)
    assert diff.max() <= 0.05 * scale, diff.max()
    ctx.close()
