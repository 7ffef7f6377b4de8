"""Batched NpT (hsmc_gpu_npt_sweeps) and its volume-move overlap scan.

* the batched scan (one kernel over the cell list the last sweep left) gives
  the same decision as the single-replica scan, which the golden vectors pin
  to the reference's check_overlap with sf (moves.c:104-112), on both of its
  paths (sweep-grid cells at least 1/sf wide, and the fallback grid);
* a replica's NpT run in a batch equals the same replica run alone, bit for
  bit (its sweeps and volume draws use only its own seed);
* the batched run samples the same density as the reference's serial NpT
  chain (npt.c:177-196) at beta P = 3.
"""
import math

import numpy as np
import pytest

from hsmc_amd import gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make(specs):
    """specs: (fcc cells per side, rho, seed, dr_max) per replica."""
    reps, confs = [], []
    for n, rho, seed, dr in specs:
        conf, box = gpu.lattice(2, n, rho)
        reps.append(dict(n=len(conf), box=box, dr_max=dr, seed=seed))
        confs.append(conf)
    ctx = gpu.Context(reps)
    for r, c in enumerate(confs):
        ctx.upload(c, r)
    return ctx


SCAN_SPECS = [(4, 0.80, 1, 0.1), (5, 0.85, 2, 0.1), (4, 0.88, 3, 0.1), (6, 0.84, 4, 0.1),
              (3, 0.80, 5, 0.1), (5, 0.86, 6, 0.1)]


def test_batched_scan_matches_single_replica_scan():
    ctx = make(SCAN_SPECS)
    ctx.sweeps(200)
    sides = [min(b / g for b, g in zip(ctx.box(r), ctx.grid(r))) for r in range(ctx.R)]
    outcomes, paths = set(), set()
    for sf in (1.0, 0.99995, 0.9999, 0.9995, 0.999, 0.998, 0.99, 0.95, 0.9):
        got = ctx.any_overlap_scaled_batch(sf)
        for r in range(ctx.R):
            assert got[r] == ctx.any_overlap_scaled([sf], r)[0], (sf, r)
            outcomes.add(int(got[r]))
            paths.add(sides[r] >= (1.0 / sf) * (1.0 + 1e-9))
    assert outcomes == {0, 1} and paths == {True, False}
    # a different factor per replica, and replicas left out (sf <= 0)
    sfs = np.array([0.999, 0.0, 0.9995, 1.0, -1.0, 0.998])
    got = ctx.any_overlap_scaled_batch(sfs)
    for r in range(ctx.R):
        want = 1 if sfs[r] <= 0 else ctx.any_overlap_scaled([sfs[r]], r)[0]
        assert got[r] == want, r
    # positions changed since the last sweep (a rescale): the cells are rebuilt
    ctx.scale(1.001, tuple(v * 1.001 for v in ctx.box(2)), r=2)
    for sf in (0.999, 0.9985):
        got = ctx.any_overlap_scaled_batch(sf)
        for r in range(ctx.R):
            assert got[r] == ctx.any_overlap_scaled([sf], r)[0], (sf, r)
    ctx.close()


@pytest.mark.parametrize("spg,sweeps", [(64, 200), (0, 530), (0, 37)])
def test_one_replica_scan_after_chained_and_rebinned_sweeps(spg, sweeps):
    """One replica: the batched scan reads the last sweep's cell list (slot
    records by sweep parity) and its grid shift (the table entry of the last
    sweep of a chained batch, or entry 0 after rebinned sweeps).  After
    chained graphs (64-sweep batches and a chained remainder), a long direct
    call (every sweep rebinned) and a short chained direct call, and after
    the sweep counter is moved, its decisions equal the single-replica scan's
    on its own grid."""
    ctx = make([(6, 0.85, 9, 0.1)])  # 864 spheres, sweep-grid side 1.0055: sf >= 0.9946 on it
    ctx.set_tuning(32, 2, spg)
    ctx.sweeps(sweeps)
    side = min(b / g for b, g in zip(ctx.box(0), ctx.grid(0)))
    outcomes = set()
    factors = (1.0, 0.9999, 0.9995, 0.999, 0.998, 0.99, 0.95)
    for sf in factors:
        got = ctx.any_overlap_scaled_batch(sf)[0]
        assert got == ctx.any_overlap_scaled([sf], 0)[0], sf
        if side >= (1.0 / sf) * (1.0 + 1e-9):  # decided on the sweep grid
            outcomes.add(int(got))
    assert outcomes == {0, 1}
    ctx.set_sweep_counter(12345)  # the last list is found by the counter's parity
    for sf in factors:
        assert ctx.any_overlap_scaled_batch(sf)[0] == ctx.any_overlap_scaled([sf], 0)[0], sf
    ctx.close()


NPT_SPECS = [(4, 0.60, 11, 0.12), (5, 0.50, 12, 0.15), (3, 0.70, 13, 0.10)]
NPT_PRESS = [3.0, 2.0, 4.0]
NPT_DV = [0.01, 0.02, 0.008]


def test_npt_batched_equals_alone():
    ctx = make(NPT_SPECS)
    counts = ctx.npt_sweeps(40, NPT_PRESS, NPT_DV)
    assert counts[:, 1].sum() > 0 and counts[:, 2].sum() + counts[:, 3].sum() > 0
    assert np.array_equal(counts[:, 0], counts[:, 1] + counts[:, 2] + counts[:, 3])
    for r, spec in enumerate(NPT_SPECS):
        alone = make([spec])
        c1 = alone.npt_sweeps(40, NPT_PRESS[r], NPT_DV[r])
        assert np.array_equal(c1[0], counts[r])
        assert alone.box(0) == ctx.box(r)
        assert np.array_equal(alone.download(0), ctx.download(r))
        alone.close()
    assert all(ctx.any_overlap_scaled_batch(1.0) == 1)
    ctx.close()


def test_npt_batched_density_matches_serial():
    """12 replicas of N = 256 at beta P = 3 against 4 runs of the serial chain;
    error bars from the run-to-run scatter of run means."""
    R, eq, stat, every = 12, 2000, 6000, 5
    ctx = make([(4, 0.63, 1 + r, 0.15) for r in range(R)])
    ctx.npt_sweeps(eq, 3.0, 0.008)
    rho = np.zeros((stat // every, R))
    vol_moves = np.zeros((R, 4), dtype=np.uint64)
    for k in range(stat // every):
        vol_moves += ctx.npt_sweeps(every, 3.0, 0.008)
        for r in range(R):
            b = ctx.box(r)
            rho[k, r] = 256 / (b[0] * b[1] * b[2])
    assert all(ctx.any_overlap_scaled_batch(1.0) == 1)
    # about one volume move per sweep: Binomial(256, 1/257) per sweep
    per_sweep = vol_moves[:, 0].astype(float) / stat
    assert abs(per_sweep.mean() - 256 / 257) < 0.05
    gpu_means = rho.mean(axis=0)
    cpu_means = []
    for seed in range(101, 105):
        o = O.System(2, 4, rho=0.63, neigh_dr=1.2, max_part=10, dr_max=0.15, seed=seed)
        o.set_params(0.15, 0.008, 3.0)
        o.sweep_npt(eq)
        r_ = []
        for _ in range(stat // every):
            o.sweep_npt(every)
            r_.append(o.box()["rho"])
        cpu_means.append(np.mean(r_))
    se_gpu = max(np.std(gpu_means, ddof=1), 0.0025) / math.sqrt(R)
    se_cpu = max(np.std(cpu_means, ddof=1), 0.0025) / math.sqrt(len(cpu_means))
    m_gpu, m_cpu = gpu_means.mean(), np.mean(cpu_means)
    assert abs(m_gpu - m_cpu) < 4.0 * math.hypot(se_gpu, se_cpu), (m_gpu, m_cpu, se_gpu, se_cpu)
    # Carnahan-Starling density at beta P = 3 is 0.6348 (N = 256 sits a little above)
    assert abs(m_gpu - 0.6348) < 0.012
    ctx.close()


# The reference's published NpT table (README.md:151-158): N = 256, packing
# fraction at each input pressure, against Brumby et al. (their values and
# standard deviations)
PUBLISHED_NPT = [(1.472, 0.249, 0.003), (2.280, 0.300, 0.002), (3.471, 0.349, 0.002),
                 (5.274, 0.399, 0.002)]


def test_npt_packing_fraction_matches_published_table():
    """8 replicas of N = 256 per pressure, all 32 in one batched context:
    the mean packing fraction agrees with the published values within three
    combined standard errors plus 0.002 for the finite size (N = 256)."""
    from hsmc_amd import analysis as A
    K, eq, stat, every = 8, 3000, 15000, 25

    def cs_density(p):
        lo, hi = 0.1, 0.99
        for _ in range(100):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if mid * A.carnahan_starling(mid)["Z"] < p else (lo, mid)
        return 0.5 * (lo + hi)

    press = [p for p, _, _ in PUBLISHED_NPT for _ in range(K)]
    dv = [{1.472: 0.03, 2.280: 0.025, 3.471: 0.02, 5.274: 0.015}[p] for p in press]
    specs = [(4, cs_density(p), 300 + r, 0.15) for r, p in enumerate(press)]
    ctx = make(specs)
    ctx.npt_sweeps(eq, press, dv)
    R = len(press)
    eta = np.zeros((stat // every, R))
    moves = np.zeros((R, 4), dtype=np.uint64)
    for k in range(stat // every):
        moves += ctx.npt_sweeps(every, press, dv)
        for r in range(R):
            b = ctx.box(r)
            eta[k, r] = math.pi / 6.0 * 256 / (b[0] * b[1] * b[2])
    assert all(ctx.any_overlap_scaled_batch(1.0) == 1)
    acc = moves[:, 1].sum() / moves[:, 0].sum()
    assert 0.03 < acc < 0.9, acc   # about 0.1 with these steps
    means = eta.mean(axis=0)
    for i, (p, ref, ref_sd) in enumerate(PUBLISHED_NPT):
        m = means[i * K:(i + 1) * K]
        se = max(np.std(m, ddof=1), 0.002) / math.sqrt(K)
        tol = 3.0 * math.hypot(se, ref_sd) + 0.002
        assert abs(m.mean() - ref) < tol, (p, m.mean(), ref, se)
    ctx.close()



def test_npt_sweeps_ideal_gas_mean_volume():
    """The batched volume move's acceptance exp(-beta P dV + (N+1) ln(V'/V))
    (moves.c:117-121, in hsmc_gpu_npt_sweeps) with a proposal symmetric in
    ln V samples V ~ Gamma(N+1, beta P) whenever no pair can overlap, so
    <V> = (N+1)/beta P.  16 replicas of 8 spheres about 11 diameters apart
    (simple cubic, tiny displacements: they never meet) must give that
    mean; a slip to N or N+2 moves it by 11%, many standard errors away."""
    N, bp, R = 8, 8e-4, 16
    rho = N * bp / (N + 1)
    reps, confs = [], []
    for r in range(R):
        conf, box = gpu.lattice(1, 2, rho)
        reps.append(dict(n=len(conf), box=box, dr_max=0.05, seed=500 + r))
        confs.append(conf)
    ctx = gpu.Context(reps)
    for r, c in enumerate(confs):
        ctx.upload(c, r)
    ctx.npt_sweeps(50, bp, 1.0)   # forget the start
    vols = []
    counts = np.zeros((R, 4), dtype=np.uint64)
    for _ in range(240):
        counts += ctx.npt_sweeps(4, bp, 1.0)
        vols.append([float(np.prod(ctx.box(r))) for r in range(R)])
    vols = np.array(vols)                       # samples x replicas
    means = vols.mean(axis=0)
    mean, err = means.mean(), means.std(ddof=1) / math.sqrt(R)
    want = (N + 1) / bp
    print(f"<V> = {mean:.1f} +- {err:.1f} (closed form {want:.1f})")
    assert abs(mean - want) < 4 * err, (mean, want, err)
    for slip in (N / bp, (N + 2) / bp):
        assert abs(mean - slip) > 8 * err, (mean, slip, err)
    assert np.all(counts[:, 0] == counts[:, 1] + counts[:, 2] + counts[:, 3])
    assert counts[:, 2].sum() == 0              # no overlap ever
    assert counts[:, 1].sum() > 0.3 * counts[:, 0].sum()


def test_c5_full_size_npt_workload():
    """The NpT workload at its own size: N = 256000 (fcc 4*40^3) at the
    Carnahan-Starling densities of beta P = 8 and 11, two replicas each in
    one batched context, ln V step 0.001.  After NpT sweeps: no pair closer
    than one diameter in any replica; N attempts per replica per sweep; every
    proposed volume move is accepted, rejected for an overlap or rejected by
    the rule; and on the N = 256000 configuration the scaled-overlap
    decisions (vol_move's scan, presst's flags) equal the oracle's."""
    specs = [(40, 0.8577, 11, 0.1), (40, 0.8577, 12, 0.1), (40, 0.9285, 13, 0.1),
             (40, 0.9285, 14, 0.1)]
    press = np.array([8.0, 8.0, 11.0, 11.0])
    ctx = make(specs)
    boxes0 = [ctx.box(r) for r in range(ctx.R)]
    ctx.reset_counters()
    sweeps = 30
    vc = ctx.npt_sweeps(sweeps, press, 0.001)
    assert np.all(vc[:, 0] == vc[:, 1] + vc[:, 2] + vc[:, 3])
    assert vc[:, 0].sum() > 0
    for r in range(ctx.R):
        c = ctx.counters(r)
        assert c["attempts"] == 256000 * sweeps, r
        assert 0 < c["accepted"] < c["attempts"]
    assert np.all(ctx.any_overlap_scaled_batch(1.0) == 1)
    # oracle comparison on a replica whose box no volume move has changed
    # (the oracle rebuilds the box from the density, as sim_box_init does)
    same = [r for r in range(ctx.R) if ctx.box(r) == boxes0[r]]
    assert same, "every replica accepted a volume move"
    r = same[-1]
    rho = specs[r][1]
    o = O.System(2, 40, rho=rho, neigh_dr=1.0, max_part=12)
    assert o.box()["lx"] == ctx.box(r)[0]
    o.set_config(ctx.download(r))
    sf = np.array([(1.0 - k * 0.0001) ** (1.0 / 3.0) for k in range(1, 21)])
    ref = o.presst_flags(sf)
    got = ctx.any_overlap_scaled(sf, r)
    assert np.array_equal(got, ref)
    print(f"replica {r} (rho {rho}): presst flags {ref.tolist()}")
    idx =np.random.default_rng(3).choice(256000, size=3000, replace=False).astype(np.int32)
    for s in (0.9999, 0.9995):
        dec = ctx.overlap_scaled(idx, np.full(len(idx), s), r)
        want = np.array([o.check_overlap(int(i), s) for i in idx], dtype=np.uint8)
        assert np.array_equal(dec, want), s
