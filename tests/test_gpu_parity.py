"""HIP kernels against the reference's golden vectors and the CPU oracle,
through the C ABI.  Integer and decision outputs must be bit-exact; FP64
positions after sweeps must be bit-exact against the oracle's restatement of
the same Philox stream."""
import math

import numpy as np
import pytest

from hsmc_amd import gpu
from oracle import oracle as O
from tests.conftest import golden_cases

pytestmark = pytest.mark.gpu

CASES = ["c1_lattice", "c1_mc", "c1_mc_dr1", "dense_sc", "n4000"]


def _ctx_for(g, name, seed=1):
    t, n, rho, ndr, mp = g[f"{name}/sys"]
    o = O.System(int(t), int(n), rho=float(rho), neigh_dr=float(ndr), max_part=int(mp))
    L = o.box()["lx"]
    conf = g[f"{name}/conf"]
    ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.1, seed=seed)])
    ctx.upload(conf)
    return ctx, o, conf, L, int(mp), float(ndr)


@pytest.mark.parametrize("name", CASES)
def test_cell_list_matches_reference(golden, name):
    ctx, o, conf, L, mp, ndr = _ctx_for(golden, name)
    n = int(math.floor(L / ndr))
    pc, nm = ctx.cell_list((n, n, n), mp)
    assert np.array_equal(pc, golden[f"{name}/ref_part_cell"])
    assert np.array_equal(nm, golden[f"{name}/ref_neigh_mat"])


@pytest.mark.parametrize("name", CASES)
def test_overlap_batch_matches_reference(golden, name):
    ctx, *_ = _ctx_for(golden, name)
    q = golden[f"{name}/overlap_q"]
    out = ctx.overlap_batch(q[:, 0].astype(np.int32), q[:, 1:])
    assert np.array_equal(out, golden[f"{name}/overlap_ref"])


@pytest.mark.parametrize("name", CASES)
def test_overlap_scaled_matches_reference(golden, name):
    ctx, o, conf, L, mp, ndr = _ctx_for(golden, name)
    q = golden[f"{name}/scaled_q"]
    ref = golden[f"{name}/scaled_ref"]
    side = L / math.floor(L / ndr)
    # the reference scans only its 27 cells; equal whenever they span 1/sf
    keep = side >= 1.0 / q[:, 1]
    out = ctx.overlap_scaled(q[keep, 0].astype(np.int32), q[keep, 1])
    assert keep.sum() > 0
    assert np.array_equal(out, ref[keep])


@pytest.mark.parametrize("name", CASES)
def test_pair_dist_matches_reference(golden, name):
    ctx, *_ = _ctx_for(golden, name)
    q = golden[f"{name}/dist_q"]
    ref = golden[f"{name}/dist_ref"]
    for sf in np.unique(q[:, 2]):
        sel = q[:, 2] == sf
        out = ctx.pair_dist(q[sel, 0].astype(np.int32), q[sel, 1].astype(np.int32), sf)
        assert np.array_equal(out, ref[sel])


@pytest.mark.parametrize("name", CASES)
def test_rdf_counts_match_reference(golden, name):
    ctx, *_ = _ctx_for(golden, name)
    h = golden[f"{name}/rdf_hist_ref"]
    rmax = float(golden[f"{name}/rdf_rmax_ref"][0])
    cnt, oor = ctx.rdf_counts(0.01, len(h), rmax)
    assert oor == 0
    assert np.array_equal(2.0 * cnt, h)


@pytest.mark.parametrize("name", [c for c in CASES if c != "c1_mc_dr1"])
def test_contact_counts_match_reference(golden, name):
    ctx, *_ = _ctx_for(golden, name)
    h = golden[f"{name}/contact_hist_ref"]
    rmax = float(golden[f"{name}/contact_rmax_ref"][0])
    cnt, oor = ctx.contact_counts(0.01, len(h), rmax)
    assert oor == 0
    assert np.array_equal(2.0 * cnt, h)


@pytest.mark.parametrize("name", CASES)
def test_presst_flags_match_reference(golden, name):
    ctx, o, conf, L, mp, ndr = _ctx_for(golden, name)
    for tag in "ab":
        dxi, xim = golden[f"{name}/presst_{tag}_args"]
        nn = int(xim / dxi)
        sf = np.array([math.pow(1 - (i + 1) * dxi, 1.0 / 3.0) for i in range(nn)])
        side = L / math.floor(L / ndr)
        if side < 1.0 / sf.min():
            continue
        out = ctx.any_overlap_scaled(sf)
        assert np.array_equal(out.astype(float), golden[f"{name}/presst_{tag}_ref"])


@pytest.mark.parametrize("name", CASES)
def test_widom_points_match_reference(golden, name):
    ctx, *_ = _ctx_for(golden, name)
    out = ctx.widom_points(golden[f"{name}/widom_q"])
    assert np.array_equal(out, golden[f"{name}/widom_ref"])


def test_widom_uniform_matches_oracle(golden):
    ctx, o, conf, L, mp, ndr = _ctx_for(golden, "c1_mc")
    o.set_config(conf)
    for sample, first, n in ((0, 0, 5000), (3, 1000, 7000), (2**33 + 5, 123456789, 4000)):
        assert ctx.widom_uniform(77, sample, first, n) == o.widom_philox(77, sample, first, n)


def test_sqrt_and_division_correctly_rounded():
    # pair distances against numpy (IEEE sqrt) on 2e5 random pairs
    rng = np.random.default_rng(3)
    n = 4000
    L = 20.0
    conf = np.zeros((n, 4))
    conf[:, 0] = np.arange(n)
    conf[:, 1:] = rng.random((n, 3)) * L
    ctx = gpu.Context([dict(n=n, box=(L, L, L))])
    ctx.upload(conf)
    i = rng.integers(0, n, 200000).astype(np.int32)
    j = rng.integers(0, n, 200000).astype(np.int32)
    out = ctx.pair_dist(i, j)
    d = conf[i, 1:] - conf[j, 1:]
    d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
    ref = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("n,rho,sweeps,seed", [(5, 0.5, 6, 3), (10, 0.85, 4, 9), (4, 0.93, 5, 21)])
def test_sweeps_match_oracle_restatement(n, rho, sweeps, seed):
    o = O.System(2, n, rho=rho, neigh_dr=1.0, max_part=10, seed=1)
    o.sweep_nvt(3)                       # a disordered start
    conf = o.config()
    L = o.box()["lx"]
    ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.1, seed=seed)])
    ctx.upload(conf)
    g = ctx.grid()
    cb = O.Checkerboard(conf, (L, L, L), g, 0.1, seed)
    ctx.sweeps(sweeps)
    cb.sweeps(sweeps)
    out = ctx.download()
    assert np.array_equal(out, cb.config())
    c = ctx.counters()
    ref = cb.counters()
    assert (c["attempts"], c["accepted"], c["left_cell"], c["sweeps"]) == tuple(int(v) for v in ref)


def test_sweeps_graph_and_direct_agree():
    o = O.System(2, 8, rho=0.8, neigh_dr=1.0, max_part=10)
    conf = o.config()
    L = o.box()["lx"]
    outs = []
    for spg in (0, 4):
        ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.12, seed=5)])
        ctx.set_tuning(32, 8, spg)
        ctx.upload(conf)
        ctx.sweeps(10)
        outs.append(ctx.download())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("lpc,cpb", [(32, 2), (64, 1)])
def test_chained_sweeps_equal_rebinned_sweeps(lpc, cpb):
    """One replica: inside a batch each sweep's last phase files the next
    sweep's cell list and only the first sweep rebins.  A direct call of
    530 sweeps (more than the sweep table holds, so every sweep rebins) must
    equal the same sweeps run as 64-sweep chained graphs plus a chained
    remainder, bit for bit, counters included."""
    o = O.System(2, 8, rho=0.85, neigh_dr=1.0, max_part=10, seed=2)
    o.sweep_nvt(2)
    conf = o.config()
    L = o.box()["lx"]
    outs = []
    for spg in (0, 64):
        ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.1, seed=7)])
        ctx.set_tuning(lpc, cpb, spg)
        ctx.upload(conf)
        ctx.sweeps(530)
        outs.append((ctx.download(), ctx.counters()))
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def test_graph_follows_replica_changes():
    """A captured sweep graph of a one-replica context carries the replica's
    record by value: after a step-size change and after a box rescale the
    graph-replayed run must equal the directly launched one, bit for bit."""
    o = O.System(2, 8, rho=0.8, neigh_dr=1.0, max_part=10)
    conf = o.config()
    L = o.box()["lx"]
    outs = []
    for spg in (0, 4):
        ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.12, seed=9)])
        ctx.set_tuning(32, 2, spg)
        ctx.upload(conf)
        ctx.sweeps(8)
        ctx.set_dr_max(0.2)
        ctx.sweeps(8)
        sf = 1.0005
        ctx.scale(sf, (L * sf, L * sf, L * sf))
        ctx.sweeps(8)
        outs.append((ctx.download(), ctx.counters()))
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def test_full_size_sweeps_keep_hard_core():
    # N = 32000, rho = 0.85 (fcc 4*20^3): 50 sweeps, no overlap, all in box
    conf, box = gpu.lattice(2, 20, 0.85)
    ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=2)])
    ctx.upload(conf)
    ctx.sweeps(50)
    c = ctx.counters()
    assert c["attempts"] == 50 * 32000 and c["sweeps"] == 50
    assert 0.3 < c["accepted"] / c["attempts"] < 0.95
    out = ctx.download()
    L = box[0]
    assert ((out[:, 1:] >= 0) & (out[:, 1:] <= L)).all()
    assert ctx.any_overlap_scaled([1.0])[0] == 1
    # contact pairs only at r >= 1: the histogram's bin of r < 1 is empty
    cnt, oor = ctx.contact_counts(0.01, 5, 1.05)
    assert oor == 0 and cnt.sum() > 0


def test_full_size_observables_match_oracle():
    """At the benchmark's own size (N = 32000, rho = 0.85, after 50 sweeps) the
    GPU observables equal the oracle's restatement of the reference on the same
    configuration, bit for bit: g(r) counts (all pairs, compute_rdf.c:110-128),
    contact counts (compute_press.c:123-165), presst flags
    (compute_press.c:243-273), pair distances and Widom hits
    (compute_widom_chem_pot.c:82-120)."""
    conf, box = gpu.lattice(2, 20, 0.85)
    ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=4)])
    ctx.upload(conf)
    ctx.sweeps(50)
    cfg = ctx.download()
    o = O.System(2, 20, rho=0.85, neigh_dr=1.05, max_part=10)
    o.set_config(cfg)
    L = box[0]
    g_cnt, g_oor = ctx.rdf_counts(0.01, 400, 5.0)
    o_cnt, o_oor = o.rdf_counts(0.01, 400, 5.0)
    assert np.array_equal(g_cnt, o_cnt) and g_oor == o_oor
    g_c, _ = ctx.contact_counts(0.01, 5, 1.05)
    o_c, _ = o.contact_counts(0.01, 5, 1.05)
    assert np.array_equal(g_c, o_c)
    sf = [(1.0 - (k + 1) * 0.0001) ** (1.0 / 3.0) for k in range(20)]
    assert np.array_equal(ctx.any_overlap_scaled(sf), o.presst_flags(sf))
    rng = np.random.default_rng(8)
    i = rng.integers(0, 32000, 2000).astype(np.int32)
    j = rng.integers(0, 32000, 2000).astype(np.int32)
    assert np.array_equal(ctx.pair_dist(i, j), [o.compute_dist(a, b) for a, b in zip(i, j)])
    pts = rng.random((200000, 3)) * L
    assert np.array_equal(ctx.widom_points(pts), o.widom_points(pts))
    ctx.close()


def test_widom_hits_match_oracle_at_n108000():
    """N = 108000 at rho = 0.90 (the learned-proposal benchmark's system): the
    decisions for 3 * 10^5 uniform trial points (almost all overlap at this
    density) and for points 1 -/+ 1e-12 from sphere centres (decided by the
    last bits of the FP64 distance) equal the reference algorithm's."""
    conf, box = gpu.lattice(2, 30, 0.90)
    ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=6)])
    ctx.upload(conf)
    ctx.sweeps(30)
    cfg = ctx.download()
    o = O.System(2, 30, rho=0.90, neigh_dr=1.0, max_part=10)
    o.set_config(cfg)
    rng = np.random.default_rng(9)
    L = box[0]
    pts = rng.random((300000, 3)) * L
    assert np.array_equal(ctx.widom_points(pts), o.widom_points(pts))
    # points at distance 1 -/+ 1e-12 from 2000 centres, kept inside [0, L)
    c = cfg[rng.integers(0, len(cfg), 2000), 1:]
    u = rng.normal(size=(2000, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    edge = np.concatenate([c + u * (1.0 - 1e-12), c + u * (1.0 + 1e-12)]) % L
    got = ctx.widom_points(edge)
    assert np.array_equal(got, o.widom_points(edge))
    assert got[:2000].sum() == 0   # inside a sphere: never insertable
    ctx.close()


def _min_image_s2(a, b, box):
    """Squared minimum-image distances (one single wrap per axis, the
    reference's compute_dist order) between rows of a and rows of b."""
    s = 0.0
    for k in range(3):
        d = a[:, None, k] - b[None, :, k]
        L = box[k]
        d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
        s = s + d * d
    return s


def test_anisotropic_box_sweeps_and_observables():
    """fcc 6 x 4 x 5 cells at rho = 0.7: three different box edges and sweep
    grid sizes.  The reference's own cell index is right only for cubes, so
    the checks are the oracle's restatement of the sweep (bit for bit, alone
    and batched beside a cubic replica) and brute force with the minimum
    image taken per axis (pair histogram on a 3 x 2 x 2 grid, Widom hits)."""
    lat, box = gpu.lattice(2, (6, 4, 5), 0.7)
    assert len(lat) == 480 and abs(box[0] * box[1] * box[2] - 480 / 0.7) < 1e-9
    g = tuple(int(L) // 2 * 2 for L in box)   # the even sweep grid, cells >= 1
    assert len(set(g)) == 3
    start = O.Checkerboard(lat, box, g, 0.4, 99)
    start.sweeps(200)                          # a melted start, with cavities
    conf = start.config()
    rep = dict(n=len(conf), box=box, dr_max=0.12, seed=17)
    ctx = gpu.Context([rep])
    ctx.upload(conf)
    assert ctx.grid() == g
    cb = O.Checkerboard(conf, box, g, 0.12, 17)
    ctx.sweeps(6)
    cb.sweeps(6)
    out = ctx.download()
    assert np.array_equal(out, cb.config())
    c = ctx.counters()
    assert (c["attempts"], c["accepted"], c["left_cell"], c["sweeps"]) == \
        tuple(int(v) for v in cb.counters())

    cube, cbox = gpu.lattice(2, 5, 0.6)
    pair = gpu.Context([dict(n=len(cube), box=cbox, dr_max=0.1, seed=3), rep])
    pair.upload(cube, 0)
    pair.upload(conf, 1)
    pair.sweeps(6)
    assert np.array_equal(pair.download(1), out)

    p = out[:, 1:]
    s = _min_image_s2(p, p, box)[np.triu_indices(len(p), 1)]
    assert s.min() >= 1.0
    dr, nb = 0.01, 200
    rmax = 1.0 + nb * dr                       # 3.0 < half the shortest edge
    assert rmax < min(box) / 2
    r = np.sqrt(s[s < rmax * rmax])
    r = r[r < rmax]
    bins = ((r - 1.0) / dr).astype(np.int64)
    cnt, oor = ctx.rdf_counts(dr, nb, rmax)
    assert np.array_equal(cnt, np.bincount(bins[bins < nb], minlength=nb))
    assert oor == int((bins >= nb).sum())

    q = np.random.default_rng(5).random((100000, 3)) * np.array(box)
    hits = ctx.widom_points(q)
    ref = np.concatenate([~(_min_image_s2(q[i:i + 10000], p, box) < 1.0).any(axis=1)
                          for i in range(0, len(q), 10000)])
    assert np.array_equal(hits.astype(bool), ref)
    assert ref.sum() > 20

    # cell-stratified insertion on the anisotropic insertion grid, each
    # point decided again by brute force over all spheres
    per_cell = 12
    h, n, grid = ctx.widom_cells(3, 1, per_cell)
    assert len(set(grid)) == 3 and n == per_cell * grid[0] * grid[1] * grid[2]
    side = np.array(box) / np.array(grid)
    pts = []
    for ix in range(grid[0]):
        for iy in range(grid[1]):
            for iz in range(grid[2]):
                c = (ix * grid[1] + iy) * grid[2] + iz
                for j in range(per_cell):
                    u = O.philox([c, j, 1, 0x57430000], [3, 0])
                    pts.append([(i + (u[k] + 0.5) * 2.0 ** -32) * side[k]
                                for k, i in enumerate((ix, iy, iz))])
    pts = np.array(pts)
    free = np.concatenate([~(_min_image_s2(pts[i:i + 10000], p, box) < 1.0).any(axis=1)
                           for i in range(0, len(pts), 10000)])
    assert h == int(free.sum()) and h > 0


@pytest.mark.parametrize("tuning", [(32, 2, 0), (64, 2, 0), (32, 1, 4)])
def test_wide_cell_single_replica_matches_oracle(tuning):
    """A dilute single replica (fcc 5^3 at rho = 0.3: sweep cells of 1.19
    diameters, so 16 slots per cell and a 192-entry stencil list) with 32-
    and 64-lane groups, direct and graph-replayed: bit for bit equal to the
    oracle's restatement of the sweep, counters included."""
    lat, box = gpu.lattice(2, 5, 0.3)
    g = tuple(int(L) // 2 * 2 for L in box)
    assert abs(box[0] / g[0] - 1.1856) < 1e-3
    start = O.Checkerboard(lat, box, g, 0.5, 7)
    start.sweeps(50)                           # a melted start
    conf = start.config()
    ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.3, seed=23)])
    ctx.set_tuning(*tuning)
    ctx.upload(conf)
    assert ctx.grid() == g
    cb = O.Checkerboard(conf, box, g, 0.3, 23)
    ctx.sweeps(12)
    cb.sweeps(12)
    assert np.array_equal(ctx.download(), cb.config())
    c = ctx.counters()
    assert (c["attempts"], c["accepted"], c["left_cell"], c["sweeps"]) == \
        tuple(int(v) for v in cb.counters())


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 2**30 + 1, 0x80000001, 0xC0000000, 0xFFFFFFFF])
def test_occupant_picks_match_oracle_and_the_rule(n):
    """The sweep kernel's occupant pick (rng_get_int's scale-and-reject rule,
    rng.c:34-36, on the Philox stream) evaluated on the device by the same
    function the phase kernel calls: bit-exact against the oracle's pick,
    always in [0, n), and uniform.  For n > 2^31 the scale is 1 and a draw
    of n or more is redrawn (half the draws at n = 2^31 + 1, a quarter at
    3 * 2^30): the redraw path runs thousands of times here, while a sweep
    cell of at most 16 occupants takes it with probability below 2^-27."""
    count = 65536
    seed, cell, sweep = 0x123456789ABCDEF, 4321, 77
    got = gpu.occupant_picks(seed, cell, sweep, n, count)
    assert got.max() < n
    ts = np.arange(0, count, 97)
    ref = np.array([O.cb_pick(seed, cell, int(t), sweep, n) for t in ts], dtype=np.uint32)
    assert np.array_equal(got[ts], ref)
    # first draws of each attempt (Philox counter (cell, t, sweep)), from the
    # oracle's KAT-pinned Philox
    key = [seed & 0xFFFFFFFF, seed >> 32]
    first = np.array([O.philox([cell, int(t), sweep, 0], key)[0] for t in ts], dtype=np.uint64)
    scale = 0xFFFFFFFF // n
    redrawn = (first // scale) >= n
    tail = (2**32 - n * scale) / 2**32   # share of draws past n whole intervals
    assert abs(redrawn.mean() - tail) < 0.05
    if tail > 0.2:
        assert redrawn.sum() > 100
    # a kept first draw is scaled down, not reduced modulo n
    assert np.array_equal(got[ts][~redrawn].astype(np.uint64), first[~redrawn] // scale)
    # uniformity: 8 equal bins of [0, n) (or the n values themselves)
    bins = min(n, 8)
    h = np.bincount((got.astype(np.float64) * bins / n).astype(np.int64), minlength=bins)
    e = count / bins
    chi2 = float(np.sum((h - e) ** 2 / e))
    assert chi2 < (bins - 1) + 6 * math.sqrt(2 * max(bins - 1, 1)) + 10, (n, h)


@pytest.mark.parametrize("n,rho,per_cell,sweeps", [(5, 0.5, 40, 100), (10, 0.8, 8, 60),
                                                   (6, 0.93, 16, 30), (30, 0.90, 12, 20)])
def test_widom_cells_match_oracle(n, rho, per_cell, sweeps):
    """Cell-stratified uniform insertion: per_cell points in every cell of
    the insertion grid (side >= 1), hits bit-exact against the oracle's
    restatement of the same Philox points (each decided by the reference's
    Widom overlap test), up to N = 108000 at rho = 0.90."""
    o = O.System(2, n, rho=rho, neigh_dr=1.0, max_part=16, seed=2)
    conf = o.config()
    L = o.box()["lx"]
    ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.1, seed=2)])
    ctx.upload(conf)
    ctx.sweeps(sweeps)
    conf = ctx.download()
    o.set_config(conf)
    hits, npts, grid = ctx.widom_cells(7, 3, per_cell)
    assert npts == per_cell * grid[0] * grid[1] * grid[2]
    assert L / grid[0] >= 1.0 and L / (grid[0] + 1) < 1.0 + 1e-9
    assert hits == o.widom_cells(7, 3, grid, per_cell)
    if rho <= 0.5:   # denser starts are still near-crystalline: no insertable volume
        assert hits > 0
    ctx.close()


def test_widom_cells_agrees_with_iid_uniform():
    """The stratified estimate and plain i.i.d. uniform insertion estimate
    the same insertion probability (rho = 0.5, N = 4000, 20 samples each)."""
    o = O.System(2, 10, rho=0.5, neigh_dr=1.0, max_part=16, seed=4)
    L = o.box()["lx"]
    ctx = gpu.Context([dict(n=4000, box=(L, L, L), dr_max=0.1, seed=4)])
    ctx.upload(o.config())
    ctx.sweeps(200)
    a, b = [], []
    for s in range(20):
        h, n, _ = ctx.widom_cells(11, s, 50)
        a.append(h / n)
        b.append(ctx.widom_uniform(13, s, 0, n) / n)
        ctx.sweeps(20)
    pa, pb = np.mean(a), np.mean(b)
    sa, sb = np.std(a, ddof=1) / math.sqrt(20), np.std(b, ddof=1) / math.sqrt(20)
    assert abs(pa - pb) < 4 * math.hypot(sa, sb), (pa, pb, sa, sb)
    ctx.close()


@pytest.mark.parametrize("n,rho,lpc,cpb,per", [(20, 0.85, 32, 2, 8), (8, 0.80, 64, 4, 5)])
def test_persistent_sweeps_equal_launched_sweeps(n, rho, lpc, cpb, per):
    """Persistent launches (one cooperative launch of `per` whole sweeps,
    grid barriers between the rebin and the phases) reproduce the
    launch-per-step sweeps bit for bit: positions, counters and sweep count,
    across launches and a partial last launch (20 sweeps)."""
    conf, box = gpu.lattice(2, n, rho)
    outs = []
    for persistent in (0, 1):
        ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=17)])
        ctx.set_tuning(lpc, cpb, per if persistent else 4, persistent)
        ctx.upload(conf)
        ctx.sweeps(20)
        outs.append((ctx.download(), ctx.counters()))
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
