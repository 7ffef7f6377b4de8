"""The CPU oracle against the reference's own outputs (tests/golden/golden.npz,
written by tests/golden/make_golden.py from fedluc/HSMC's translation units)
and against known-answer tests and brute force."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden_cases


def _system(g, name):
    t, n, rho, ndr, mp = g[f"{name}/sys"]
    return O.System(int(t), int(n), rho=float(rho), neigh_dr=float(ndr), max_part=int(mp),
                    dr_max=0.1, seed=1)


def test_mt19937_known_answer():
    # C++11 [rand.predef]: 10000th output of mt19937 with the default seed 5489
    m = O.MT(5489)
    for _ in range(9999):
        m.next()
    assert m.next() == 4123659995


def _numpy_mt_stream(seed, n):
    """Raw 32-bit outputs of numpy's own MT19937 (legacy RandomState, seeded
    by init_genrand(seed) as gsl_rng_mt19937 seeds it): an implementation
    independent of the oracle's."""
    return np.random.RandomState(seed).randint(0, 2**32, size=n, dtype=np.uint64)


def test_mt_stream_matches_an_independent_mt19937():
    for seed in (1, 7, 4357, 123456789):
        ref = _numpy_mt_stream(seed, 3000)
        m = O.MT(seed)
        assert np.array_equal(np.array([m.next() for _ in range(3000)], dtype=np.uint64), ref)


def test_mt_double_is_draw_over_max():
    # rng_get_double = gsl_rng_get / gsl_rng_max (rng.c:29-31), max = 2^32 - 1
    # for mt19937: the closed interval [0, 1]
    ref = _numpy_mt_stream(7, 2000).astype(np.float64) / 4294967295.0
    m = O.MT(7)
    assert np.array_equal(np.array([m.double() for _ in range(2000)]), ref)


def test_mt_int_discards_the_tail_beyond_whole_intervals():
    """rng_get_int is gsl_rng_uniform_int (rng.c:34-36; its argument is an
    int, so n < 2^31), documented to reach [0, n) "by scaling down and/or
    discarding samples": a draw is divided by the integer scale
    floor((2^32-1)/n), and a quotient of n or more (the tail past n whole
    intervals) is discarded and drawn again.  With n = 2^30 + 1 (scale 3) a
    quarter of all draws fall in that tail, with n = 0x55555556 (scale 2) a
    third, so the outputs must be exactly the in-range quotients of the
    independent raw stream, in order.  A modulo reduction, a scale of 2^32/n,
    a multiply-shift map or a tail folded back into range all fail."""
    for n, seed in ((2**30 + 1, 11), (0x55555556, 12)):
        scale = 0xFFFFFFFF // n
        raw = _numpy_mt_stream(seed, 4000)
        q = raw // scale
        kept = q[q < n]
        assert 0.2 < 1.0 - len(kept) / len(raw) < 0.4   # the discard path is taken often
        m = O.MT(seed)
        got = np.array([m.int(n) for _ in range(len(kept))], dtype=np.uint64)
        assert np.array_equal(got, kept)
        # each discarded draw consumed exactly one output: the stream now
        # stands right after the last kept draw
        last = int(np.nonzero(q < n)[0][-1])
        assert m.next() == int(_numpy_mt_stream(seed, last + 2)[last + 1])


def test_mt_int_scales_down_into_equal_intervals():
    """For small n the rule scales: the output is a nondecreasing function of
    the raw draw, each value taking one interval of equal width (up to one
    unit), and all n values appear equally often (chi-square).  A modulo or
    parity rule is not monotone in the draw."""
    for n, seed in ((2, 21), (7, 22), (500, 23)):
        raw = _numpy_mt_stream(seed, 60000)
        m = O.MT(seed)
        k = np.array([m.int(n) for _ in range(60000)], dtype=np.int64)
        assert k.min() >= 0 and k.max() < n
        # no draw of these streams falls in the discarded tail (at most n-1
        # values out of 2^32), so outputs and draws stay aligned
        order = np.argsort(raw, kind="stable")
        assert np.all(np.diff(k[order]) >= 0)
        lo = np.array([raw[k == v].min() for v in range(n) if np.any(k == v)], dtype=np.float64)
        width = 2.0**32 / n
        if n <= 7:
            assert np.all(np.abs(lo - np.arange(n) * width) < 2e-3 * width)
        counts = np.bincount(k, minlength=n)
        e = len(k) / n
        chi2 = float(np.sum((counts - e) ** 2 / e))
        assert chi2 < (n - 1) + 6 * math.sqrt(2 * (n - 1)) + 10, (n, chi2)


def test_vol_move_ideal_gas_mean_volume():
    """vol_move's acceptance exp(-beta P dV + (N+1) ln(V'/V)) (moves.c:117-121)
    with a proposal symmetric in ln V has the stationary density
    V^(N+1) exp(-beta P V) in ln V, i.e. V ~ Gamma(N+1, beta P) when no pair
    ever overlaps: <V> = (N+1)/beta P.  Eight spheres about 11 diameters
    apart that barely move never overlap, so the oracle's serial NpT chain
    must give that mean; a slip to N or N+2 moves it by 11%, far outside the
    blocking error of the run."""
    from hsmc_amd import analysis as A
    N, bp = 8, 8e-4
    # 3 cells per side (the reference keeps its initial cell count in NpT,
    # so the box may not shrink below 3 diameters; Gamma(9) never gets there)
    o = O.System(1, 2, rho=N * bp / (N + 1), neigh_dr=7.0, max_part=10, seed=5)
    assert o.ncells() == 3
    o.set_params(dr_max=0.05, dv_max=1.0, press=bp)
    o.sweep_npt(500)   # forget the start
    vols = np.empty(24000)
    for k in range(len(vols)):
        o.sweep_npt(1)
        vols[k] = o.box()["vol"]
    mean = vols.mean()
    err = A.blocking_error(vols)
    assert err > 0
    want = (N + 1) / bp
    assert abs(mean - want) < 4 * err, (mean, want, err)
    for slip in (N / bp, (N + 2) / bp):
        assert abs(mean - slip) > 8 * err, (mean, slip, err)
    # the volume chain moved: several hundred accepted moves per thousand
    c = o.counters()
    assert c[4] > 0.3 * c[3]


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_lattice_closed_form():
    # fcc at rho = 0.5: a = (4/0.5)^(1/3) = 2 exactly (sim_info.c:51-52)
    s = O.System(2, 5, rho=0.5, neigh_dr=1.05, max_part=10)
    b = s.box()
    assert b["lx"] == 10.0 and b["vol"] == 1000.0
    c = s.config()
    assert c.shape == (500, 4)
    assert np.array_equal(c[:4, 1:], [[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1]])
    # nearest-neighbour distance a/sqrt(2): no pair closer
    d = c[:, None, 1:] - c[None, :, 1:]
    d -= 10.0 * np.round(d / 10.0)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(500) * 99
    assert abs(r.min() - math.sqrt(2.0)) < 1e-12


@pytest.mark.parametrize("name", ["c1_lattice", "c1_mc", "c1_mc_dr1", "dense_sc", "n4000"])
def test_oracle_matches_reference(golden, name):
    g = golden
    o = _system(g, name)
    b = g[f"{name}/ref_box"]
    box = o.box()
    assert b[0] == box["vol"] and b[1] == box["lx"] and b[2] == box["ly"] and b[3] == box["lz"]
    if f"{name}/ref_lattice" in g.files:
        assert np.array_equal(o.config(), g[f"{name}/ref_lattice"])
    o.set_config(g[f"{name}/conf"])
    pc, nm = o.cell_list()
    assert np.array_equal(pc, g[f"{name}/ref_part_cell"])
    assert np.array_equal(nm, g[f"{name}/ref_neigh_mat"])
    q = g[f"{name}/overlap_q"]
    dec = [o.overlap_at(int(r[0]), r[1], r[2], r[3]) for r in q]
    assert np.array_equal(np.array(dec, dtype=np.uint8), g[f"{name}/overlap_ref"])
    q = g[f"{name}/scaled_q"]
    dec = [o.check_overlap(int(r[0]), r[1]) for r in q]
    assert np.array_equal(np.array(dec, dtype=np.uint8), g[f"{name}/scaled_ref"])
    q = g[f"{name}/dist_q"]
    d = np.array([o.compute_dist(int(r[0]), int(r[1]), r[2]) for r in q])
    assert np.array_equal(d, g[f"{name}/dist_ref"])
    h = g[f"{name}/rdf_hist_ref"]
    cnt, oor = o.rdf_counts(0.01, len(h), float(g[f"{name}/rdf_rmax_ref"][0]))
    assert oor == 0 and np.array_equal(2.0 * cnt, h)
    if f"{name}/contact_hist_ref" in g.files:
        h = g[f"{name}/contact_hist_ref"]
        cnt, oor = o.contact_counts(0.01, len(h), float(g[f"{name}/contact_rmax_ref"][0]))
        assert oor == 0 and np.array_equal(2.0 * cnt, h)
    for tag in "ab":
        dxi, xim = g[f"{name}/presst_{tag}_args"]
        nn = int(xim / dxi)
        sf = [math.pow(1 - (i + 1) * dxi, 1.0 / 3.0) for i in range(nn)]
        assert np.array_equal(o.presst_flags(sf).astype(float), g[f"{name}/presst_{tag}_ref"])
    assert np.array_equal(o.widom_points(g[f"{name}/widom_q"]), g[f"{name}/widom_ref"])
    mv = g[f"{name}/moves_q"]
    acc = [o.part_move_scripted(int(r[0]), r[1], r[2], r[3]) for r in mv]
    assert np.array_equal(np.array(acc, dtype=np.uint8), g[f"{name}/moves_acc_ref"])
    assert np.array_equal(o.config(), g[f"{name}/moves_conf_ref"])
    assert np.array_equal(o.cell_list()[0], g[f"{name}/moves_part_cell_ref"])


def test_overlap_against_brute_force(golden):
    # check_overlap through the cell list equals an all-particle scan
    # (the variant kept commented out at moves.c:196-207)
    g = golden
    o = _system(g, "c1_mc")
    conf = g["c1_mc/conf"]
    o.set_config(conf)
    L = o.box()["lx"]
    q = g["c1_mc/overlap_q"][:400]
    for r in q:
        i = int(r[0])
        d = r[1:] - conf[:, 1:]
        d -= L * np.where(d > L / 2, 1, np.where(d < -L / 2, -1, 0))
        s = (d ** 2).sum(1)
        s[i] = 9.0
        assert o.overlap_at(i, r[1], r[2], r[3]) == int((np.sqrt(s) < 1.0).any())


def test_rdf_against_numpy(golden):
    g = golden
    o = _system(g, "c1_mc")
    conf = g["c1_mc/conf"]
    o.set_config(conf)
    L = 10.0
    d = conf[:, None, 1:] - conf[None, :, 1:]
    d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
    r = np.sqrt(((d * d).sum(-1)))[np.triu_indices(len(conf), 1)]
    rmax = 0.01 * 400 + 1.0
    sel = r[r < rmax]
    bins = ((sel - 1.0) / 0.01).astype(int)
    ref = np.bincount(bins, minlength=400)
    cnt, oor = o.rdf_counts(0.01, 400, rmax)
    assert np.array_equal(cnt, ref) and oor == 0


def test_checkerboard_sweep_invariants():
    conf = O.System(2, 6, rho=0.85, neigh_dr=1.0, max_part=10).config()
    L = 6 * (4 / 0.85) ** (1 / 3)
    cb = O.Checkerboard(conf, (L, L, L), (8, 8, 8), 0.1, seed=5)
    cb.sweeps(20)
    c = cb.config()
    cnt = cb.counters()
    assert cnt[0] == 20 * len(conf) and cnt[3] == 20
    assert 0 < cnt[1] < cnt[0]
    d = c[:, None, 1:] - c[None, :, 1:]
    d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
    s = (d * d).sum(-1) + np.eye(len(c)) * 9
    assert s.min() >= 1.0
    assert ((c[:, 1:] >= 0) & (c[:, 1:] <= L)).all()


def test_checkerboard_randoms_are_a_permutation():
    conf = O.System(2, 4, rho=0.5, neigh_dr=1.0, max_part=10).config()
    cb = O.Checkerboard(conf, (8.0, 8.0, 8.0), (6, 6, 6), 0.1, seed=11)
    seen = set()
    for s in range(200):
        sh, perm = cb.sweep_randoms(s)
        assert sorted(perm.tolist()) == list(range(8))
        assert ((sh >= 0) & (sh < 8.0 / 6)).all()
        seen.add(tuple(perm.tolist()))
    assert len(seen) > 150
