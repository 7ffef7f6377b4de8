"""Headline benchmark: MC move attempts/s of the checkerboard sweep at
N = 32000, rho = 0.85 (fcc 4*20^3, dr_max = 0.1), on one MI355X, and the
GPU seconds to sigma(beta mu_ex) = 0.01 at rho = 0.90 (N = 108000), learned
against uniform Widom insertion.

A "step" is one sweep (N attempted moves).  The system is melted from the
fcc lattice by a stated number of sweeps before the autotuner and the
warm-up run, so the timed sweeps see the equilibrium fluid.  Inputs are
resident in HBM when the timed region starts.  Prints one JSON line: metric,
value, unit, timing, the roofline of the sublattice-phase kernel, the Widom
comparison (a `widom` object) and the 1-core CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per attempted move of the sweep: per sweep N*(24+24)
# FP64 positions read+written, N*(4+4) cell-slot ids, C*(4+4) cell counts;
# at N = 32000 on a 32^3 grid 2,054,144 B / 32000 = 64.19 B per attempt
N_PART = 32000
GRID_CELLS = 32 ** 3
# sweeps that melt the fcc start before tuning and timing (the fluid's
# acceptance, about 0.53 at dr_max 0.1, is reached well within them)
SETUP_SWEEPS = 5000
BYTES_PER_SWEEP = N_PART * 48 + N_PART * 8 + GRID_CELLS * 8
BYTES_PER_ATTEMPT = BYTES_PER_SWEEP / N_PART


def _cpu_replica_rate(args) -> float:
    """Attempts/s of the oracle's serial sweep on one N = 32000 replica."""
    seed, target_s = args
    from oracle import oracle as O
    o = O.System(2, 20, rho=0.85, neigh_dr=1.05, max_part=10, dr_max=0.1, seed=seed)
    w, _ = o.time_sweeps(5)
    n = max(5, int(target_s / max(w / 5, 1e-6)))
    w, _ = o.time_sweeps(n)
    return N_PART * n / w


def cpu_baseline(target_s: float = 12.0) -> dict:
    """The oracle's serial reference-algorithm sweep (1 core) on the same
    system: reference grid neigh_list 1.05 (31^3 cells), max_part 10.  Also
    the all-cores aggregate: one independent replica per worker process, at
    most 16 (a GPU's share of the host)."""
    import multiprocessing as mp

    from oracle import oracle as O
    o = O.System(2, 20, rho=0.85, neigh_dr=1.05, max_part=10, dr_max=0.1, seed=1)
    w, _ = o.time_sweeps(10)
    n = max(10, int(target_s / max(w / 10, 1e-6)))
    w, c = o.time_sweeps(n)
    rate = N_PART * n / w
    workers = max(1, min(16, os.cpu_count() or 1))
    with mp.get_context("fork").Pool(workers) as pool:
        rates = pool.map(_cpu_replica_rate, [(2 + k, 8.0) for k in range(workers)])
    return {"value": rate, "unit": "MC move attempts/s", "cores": 1, "kind": "port",
            "sample": f"{n} serial sweeps ({n * N_PART} part_move attempts) of the same "
                      f"N=32000 rho=0.85 system from the fcc lattice, oracle/liborc.so, "
                      f"gcc -O2, {w:.1f} s wall",
            "all_cores": {"value": float(sum(rates)), "cores": workers,
                          "sample": f"{workers} independent replicas, one per process, "
                                    f"about 8 s of serial sweeps each"}}


def observables(ctx, no_cpu: bool) -> dict:
    """Per-sample cost of the observables on the same N = 32000 system, each
    sample right after a sweep (so it pays its own neighbour-grid build):
    g(r) 0.01 sigma to 5 sigma (400 bins), the contact histogram (5 bins of
    0.01), and 10^7 uniform Widom insertions.  Beside them, one g(r) sample
    of the oracle's restatement of compute_rdf (all pairs, compute_rdf.c:118-126)
    on the same configuration, on one core."""
    def per_sample(fn, reps):
        total = 0.0
        for _ in range(reps):
            ctx.sweeps(1)
            ctx.sync()
            t0 = time.perf_counter()
            fn()
            total += time.perf_counter() - t0
        return total / reps
    out = {
        "rdf_400_bins_ms": 1e3 * per_sample(lambda: ctx.rdf_counts(0.01, 400, 5.0), 20),
        "contact_5_bins_ms": 1e3 * per_sample(lambda: ctx.contact_counts(0.01, 5, 1.05), 20),
        "widom_1e7_uniform_ms": 1e3 * per_sample(lambda: ctx.widom_uniform(7, 0, 0, 10_000_000), 10),
    }
    if not no_cpu:
        from oracle import oracle as O
        o = O.System(2, 20, rho=0.85, neigh_dr=1.05, max_part=10, dr_max=0.1, seed=1)
        o.set_config(ctx.download())
        t0 = time.perf_counter()
        o.rdf_counts(0.01, 400, 5.0)
        out["cpu_rdf_400_bins_ms"] = 1e3 * (time.perf_counter() - t0)
    return out


def traffic_from_profile(tuning: dict):
    """HBM bytes per sublattice-phase launch at this launch shape (lanes per
    cell, cells per workgroup; graph batching does not change a dispatch),
    from the calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE runs in profiles/
    (None when that shape has not been profiled)."""
    path = os.path.join(ROOT, "profiles", "sweep_phase_n32000_pmc.json")
    if not os.path.exists(path):
        return None
    key = f"{tuning['lanes_per_cell']},{tuning['cells_per_block']}"
    try:
        with open(path) as f:
            entry = json.load(f)["by_tuning"].get(key)
        return None if entry is None else float(entry["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def calibrate_copy(mib: int) -> None:
    """One device-to-device copy of `mib` MiB through the HIP runtime alone
    (ctypes; neither torch nor the library is loaded): the known-size
    reference for the rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(mib << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    for name, ptr in (("src", src), ("dst", dst)):
        if hip.hipMalloc(C.byref(ptr), nbytes) != 0:
            raise SystemExit(f"hipMalloc of the {name} buffer failed")
    if hip.hipMemset(src, 1, nbytes) != 0 or hip.hipDeviceSynchronize() != 0:
        raise SystemExit("hipMemset failed")
    if hip.hipMemcpy(dst, src, nbytes, 3) != 0 or hip.hipDeviceSynchronize() != 0:  # 3: device to device
        raise SystemExit("hipMemcpy failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"[bench] copied {mib} MiB device to device", file=sys.stderr)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed sweeps (default: 20000 for n32000, 2000 for batched, 20 for npt)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed sweeps first (default: 1000, 200, 5 for the same workloads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["n32000", "batched", "widom", "npt", "widom_npt"],
                    default="n32000",
                    help="n32000: the headline single system; batched: 256 replicas "
                         "N = 4 n^3, n in 3..15, rho in [0.30, 0.93]; widom: GPU seconds to "
                         "sigma(beta mu_ex) = 0.01 at N = 108000, rho = 0.90, learned vs uniform; "
                         "npt: 64 NpT replicas of N = 256000 at beta P in {5, 8, 11}; "
                         "widom_npt: learned vs uniform Widom on NpT configurations of N = 256000 "
                         "at beta P in {5, 8, 11} (--warmup NVT melt sweeps, --steps NpT sweeps)")
    ap.add_argument("--sigma", type=float, default=0.01, help="widom: target sigma(beta mu_ex)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", type=str, default="",
                    help="lanes_per_cell,cells_per_block,sweeps_per_graph (skips the autotuner)")
    ap.add_argument("--no-phase-timing", action="store_true",
                    help="skip the per-launch HIP-event timing pass (for profiler runs)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (HIP runtime, no "
                         "torch) and exit: the known-size reference for rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--no-widom", action="store_true",
                    help="leave the Widom comparison out of the headline line")
    ap.add_argument("--widom-configs", type=int, default=128,
                    help="configurations sampled along the Widom chain")
    ap.add_argument("--widom-gap", type=int, default=100, help="sweeps between Widom samples")
    ap.add_argument("--profile-widom", action="store_true",
                    help="only run the Widom kernels on the rho = 0.90 system (for rocprofv3)")
    ap.add_argument("--state", type=str, default="",
                    help="melted configuration (.npy): loaded in place of the melt when the file "
                         "exists, else written after the melt (profiled runs load one, so the "
                         "profiler sees only the timed sweeps)")
    ap.add_argument("--save-only", action="store_true",
                    help="with --state: melt, write the file and exit")
    ap.add_argument("--no-observables", action="store_true",
                    help="leave the per-sample observable timings out")
    args = ap.parse_args()
    defaults = {"n32000": (20000, 1000), "batched": (2000, 200), "widom": (0, 0), "npt": (20, 5),
                "widom_npt": (2000, 20000)}
    if args.steps is None:
        args.steps = defaults[args.workload][0]
    if args.warmup is None:
        args.warmup = defaults[args.workload][1]
    if args.calibrate_copy_mib:
        return calibrate_copy(args.calibrate_copy_mib)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent replicas shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        return batched(args)
    if args.workload == "widom":
        return widom(args)
    if args.workload == "npt":
        return npt(args)
    if args.workload == "widom_npt":
        return widom_npt(args)
    if args.profile_widom:
        return profile_widom(args)

    from hsmc_amd import gpu
    conf, box = gpu.lattice(2, 20, 0.85)
    ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=1)])
    ctx.upload(conf)
    print(f"[bench] grid {ctx.grid()}  N={len(conf)}", file=sys.stderr, flush=True)
    if not _load_state(ctx, args.state):
        ctx.sweeps(SETUP_SWEEPS)   # melt the lattice; the autotuner restores this state
        ctx.sync()
        if _save_state(ctx, args):
            return
    if args.tuning:
        ctx.set_tuning(*(int(v) for v in args.tuning.split(",")))
        tuning = ctx.get_tuning()
    else:
        tuning = ctx.get_tuning() if args.no_autotune else ctx.autotune(512)
    print(f"[bench] melted by {SETUP_SWEEPS} sweeps; tuning {tuning}", file=sys.stderr, flush=True)

    ctx.sweeps(args.warmup)
    ctx.sync()
    ctx.reset_counters()
    t0 = time.perf_counter()
    ms = ctx.time_sweeps(args.steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    c = ctx.counters()
    assert c["attempts"] == args.steps * N_PART, c
    secs = ms / 1e3
    value = N_PART * args.steps / secs
    print(f"[bench] {args.steps} sweeps in {secs:.3f} s (events), {wall:.3f} s (wall); "
          f"acceptance {c['accepted'] / c['attempts']:.4f}", file=sys.stderr, flush=True)

    # dominant kernel: the sublattice-phase kernel, timed per launch
    avg_ms, launches = (float("nan"), 0) if args.no_phase_timing else ctx.time_phase_kernel(200)
    per_launch_bytes = BYTES_PER_ATTEMPT * N_PART / 8.0
    achieved = per_launch_bytes / (avg_ms / 1e3) / 1e9
    if ctx.any_overlap_scaled([1.0])[0] != 1:
        raise SystemExit("overlap found after the benchmark sweeps")

    line = {
        "metric": "MC move attempts/s (N=32k, rho=0.85)",
        "value": value,
        "unit": "MC move attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (fcc 4*20^3 lattice start melted by setup_sweeps, Philox proposals)",
        "config": {"workload": "NVT hard spheres N=32000 (fcc 4*20^3) rho=0.85 dr_max=0.1, "
                               "checkerboard sweep on a 32^3 grid",
                   "n_part": N_PART, "rho": 0.85, "dr_max": 0.1, "grid": list(ctx.grid()),
                   "setup_sweeps": SETUP_SWEEPS, "tuning": tuning,
                   "timed_launches": "graph replays only (a graph of exactly the remaining "
                                     "steps mod sweeps_per_graph is captured before timing)",
                   "acceptance": c["accepted"] / c["attempts"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(tuning),
                     "kernel": "k_phase (sublattice phase)", "avg_launch_ms": avg_ms,
                     "launches_timed": launches,
                     "algorithmic_bytes_per_launch": per_launch_bytes},
    }
    if not args.no_observables:
        line["observables"] = observables(ctx, args.no_cpu)
    ctx.close()
    if not args.no_widom:
        line["widom"] = widom_headline(args)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


def batched(args) -> None:
    """256 independent replicas swept together (one graph launch per batch of
    sweeps); prints its own JSON line."""
    import numpy as np

    from hsmc_amd import gpu
    rng = np.random.default_rng(2026)
    reps = []
    for r in range(256):
        n = int(rng.integers(3, 16))
        rho = float(rng.uniform(0.30, 0.93))
        conf, box = gpu.lattice(2, n, rho)
        reps.append(dict(n=len(conf), box=box, dr_max=0.1, seed=1 + r, conf=conf))
    total = sum(r["n"] for r in reps)
    ctx = gpu.Context(reps)
    for r, rep in enumerate(reps):
        ctx.upload(rep["conf"], r)
    tuning = ctx.autotune(32) if not args.tuning else None
    if args.tuning:
        ctx.set_tuning(*(int(v) for v in args.tuning.split(",")))
        tuning = ctx.get_tuning()
    print(f"[bench] batched: {len(reps)} replicas, {total} spheres, tuning {tuning}",
          file=sys.stderr, flush=True)
    ctx.sweeps(args.warmup)
    ctx.sync()
    steps = args.steps
    ms = ctx.time_sweeps(steps)
    value = total * steps / (ms / 1e3)
    acc = [ctx.counters(r) for r in range(len(reps))]
    print(json.dumps({
        "metric": "MC move attempts/s (256 batched replicas, N=108..13500, rho=0.30..0.93)",
        "value": value, "unit": "MC move attempts/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": ms / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (fcc lattice starts, Philox proposals)",
        "config": {"workload": "256 NVT replicas, N = 4 n^3 with n uniform in 3..15, rho uniform "
                               "in [0.30, 0.93], one Philox seed each", "total_spheres": total,
                   "tuning": tuning,
                   "acceptance": sum(a["accepted"] for a in acc) / max(1, sum(a["attempts"] for a in acc))},
    }))


def npt(args) -> None:
    """64 NpT replicas of N = 256000 (fcc 4*40^3) at beta P = 5, 8, 11 (21, 21
    and 22 replicas, each started at the Carnahan-Starling density of its
    pressure), ln-volume step 0.001: one step is one NpT sweep of every
    replica (a checkerboard sweep, then each replica's Binomial(N, 1/(N+1))
    volume moves, each checked by an all-pairs scaled-overlap reduction)."""
    import numpy as np

    from hsmc_amd import gpu

    levels = [5.0, 8.0, 11.0]
    press = [levels[0]] * 21 + [levels[1]] * 21 + [levels[2]] * 22
    R, n_cells, dv = len(press), 40, 0.001
    reps, confs = [], []
    for r, p in enumerate(press):
        conf, box = gpu.lattice(2, n_cells, _cs_density(p))
        reps.append(dict(n=len(conf), box=box, dr_max=0.1, seed=1 + r))
        confs.append(conf)
    total = sum(r["n"] for r in reps)
    ctx = gpu.Context(reps)
    if args.tuning:
        ctx.set_tuning(*(int(v) for v in args.tuning.split(",")))
    for r, c in enumerate(confs):
        ctx.upload(c, r)
    del confs
    if not args.tuning and not args.no_autotune:
        ctx.autotune(16)  # NVT trials of the same phase kernels; positions are restored
    print(f"[bench] npt: {R} replicas, {total} spheres, tuning {ctx.get_tuning()}",
          file=sys.stderr, flush=True)
    ctx.npt_sweeps(args.warmup, press, dv)
    ctx.sync()
    ctx.reset_counters()
    t0 = time.perf_counter()
    vol = ctx.npt_sweeps(args.steps, press, dv)
    ctx.sync()
    secs = time.perf_counter() - t0
    acc = [ctx.counters(r) for r in range(R)]
    attempts = sum(a["attempts"] for a in acc)
    assert attempts == total * args.steps, (attempts, total * args.steps)
    if not all(ctx.any_overlap_scaled_batch(1.0) == 1):
        raise SystemExit("overlap found after the NpT sweeps")
    dens = {}
    for p in levels:
        rs = [r for r in range(R) if press[r] == p]
        vals = [reps[r]["n"] / float(np.prod(ctx.box(r))) for r in rs]
        dens[str(p)] = {"mean_density": float(np.mean(vals)), "carnahan_starling": _cs_density(p)}
    print(json.dumps({
        "metric": "MC move attempts/s (64 NpT replicas, N=256000, beta P in {5,8,11})",
        "value": attempts / secs, "unit": "MC move attempts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (fcc 4*40^3 starts at the Carnahan-Starling density, Philox proposals)",
        "config": {"workload": "64 NpT replicas (21 at beta P = 5, 21 at 8, 22 at 11), N = 256000 "
                               "each, max ln-volume step 0.001, dr_max 0.1",
                   "total_spheres": total, "tuning": ctx.get_tuning(),
                   "acceptance": sum(a["accepted"] for a in acc) / attempts},
        "volume_moves": {"proposed": int(vol[:, 0].sum()), "accepted": int(vol[:, 1].sum()),
                         "rejected_overlap": int(vol[:, 2].sum()),
                         "rejected_rule": int(vol[:, 3].sum()),
                         "per_second": float(vol[:, 0].sum()) / secs},
        "density_after_run": dens,
    }))


def _timed(fn):
    t0 = time.perf_counter()
    out = fn()
    return out, time.perf_counter() - t0


def _widom_chain(ctx, advance, target, S, gap, n_u, n_l, seed=0, train_points=1_000_000):
    """Uniform against learned Widom along one running chain of `ctx`
    (replica 0).  The network is trained on the current configuration first
    (timed, and charged to the learned method).  Then S samples, each `gap`
    chain steps after the last (advance(gap); reported apart and charged to
    neither method, since both use the same configurations).  At each sample
    every method runs at several insertion counts on the same configuration:
    uniform (cell-stratified) at about 4 n_u, n_u, n_u/4 and n_u/16 points,
    learned at 4 n_l, n_l and n_l/4 insertions, each learned call scoring
    the configuration again and rebuilding its cumulative mass (charged).  Per count,
    sigma(beta mu_ex) comes from blocking over the per-configuration
    estimates (hsmc_stat.py:6-50), so configurational variance is sampled;
    GPU seconds to the target sigma follow the 1/samples law from the
    measured seconds.  Each method is reported at its fastest count (for
    the ratio counting the chain's sweeps, at its fastest count by that
    measure)."""
    import math

    import numpy as np

    from hsmc_amd import analysis as A
    from hsmc_amd import gpu
    voxel, mix = 0.25, 0.9
    net = gpu.Mlp(ctx, seed=5)
    ctx.sync()
    loss, t_train = _timed(lambda: net.train(voxel, 777 + seed, 0, train_points, epochs=10,
                                             batch=4096, lr=2e-3, beta=20.0))
    # uniform insertion in its fast, cell-stratified form: per_cell points in
    # each cell of the insertion grid
    _, _, grid = ctx.widom_cells(0, 0, 0)
    cells = grid[0] * grid[1] * grid[2]
    per_cell = max(1, round(n_u / cells))
    pcs = sorted({max(1, round(per_cell * f)) for f in (4, 1, 1 / 4, 1 / 16)}, reverse=True)
    nls = sorted({max(1, round(n_l * f)) for f in (4, 1, 1 / 4)}, reverse=True)
    # the i.i.d. uniform kernel (widom_insertion's point stream) on the
    # starting configuration, for its insertion rate
    _, t_iid = _timed(lambda: ctx.widom_uniform(99, 0, 0, pcs[0] * cells))
    pu = {pc: [] for pc in pcs}
    tu = {pc: [] for pc in pcs}
    hu = {pc: 0 for pc in pcs}
    pl = {n: [] for n in nls}
    tl = {n: [] for n in nls}
    hl = {n: 0 for n in nls}
    t_chain = 0.0
    for k in range(S):
        _, dt = _timed(lambda: (advance(gap), ctx.sync()))
        t_chain += dt
        for pc in pcs:
            (h, _, _), dt = _timed(lambda: ctx.widom_cells(1234 + seed, k, pc))
            pu[pc].append(h / (pc * cells))
            tu[pc].append(dt)
            hu[pc] += h
        for n in nls:
            net.forget()
            r, dt = _timed(lambda: net.widom(voxel, mix, 4321 + seed, k, n))
            pl[n].append(r["estimate"])
            tl[n].append(dt)
            hl[n] += r["hits"]
    # the variance per insertion on one fixed configuration (the last): 8
    # repeated estimates per method at its largest count, so configurational
    # variance is absent
    n_u0 = pcs[0] * cells
    rep_u = [ctx.widom_cells(777 + k, 0, pcs[0])[0] / n_u0 for k in range(8)]
    rep_l = [net.widom(voxel, mix, 888 + k, 0, nls[0])["estimate"] for k in range(8)]
    net.close()

    def summary(p, t, n_per, fixed, hits):
        p = np.asarray(p, dtype=np.float64)
        mean = float(p.mean())
        sig = A.blocking_error(p) / mean if mean > 0 else float("inf")
        secs = float(sum(t))
        factor = (sig / target) ** 2
        return dict(beta_mu=-math.log(mean) if mean > 0 else float("inf"), sigma=sig,
                    p_insert=mean, configurations=S, insertions=n_per * S, hits=int(hits),
                    seconds_measured=fixed + secs,
                    seconds_to_target=fixed + secs * factor,
                    seconds_to_target_with_chain=fixed + (secs + t_chain) * factor,
                    blocking_relative=(A.blocking_std(p)[:, 0] / mean).tolist() if mean > 0 else [],
                    ms_per_configuration=1e3 * secs / S, insertions_per_s=n_per * S / secs)

    def best(table, key="seconds_to_target"):
        return min(table, key=lambda d: d[key])

    unis = []
    for pc in pcs:
        d = summary(pu[pc], tu[pc], pc * cells, 0.0, hu[pc])
        d["form"] = f"cell-stratified, {pc} points in each of {cells} cells"
        unis.append(d)
    lrns = []
    for n in nls:
        d = summary(pl[n], tl[n], n, t_train, hl[n])
        d["training_seconds"] = t_train
        d["training_loss"] = loss
        lrns.append(d)
    uni, lrn = best(unis), best(lrns)
    uni["iid_kernel_insertions_per_s"] = n_u0 / t_iid
    var_u = float(np.var(rep_u, ddof=1)) * n_u0
    var_l = float(np.var(rep_l, ddof=1)) * nls[0]
    brief = lambda d: {k: d[k] for k in ("insertions", "sigma", "ms_per_configuration",
                                         "seconds_to_target", "seconds_to_target_with_chain")}
    # counting the chain's sweeps too, each method's fastest count can differ
    wc = "seconds_to_target_with_chain"
    return {
        "ratio": uni["seconds_to_target"] / lrn["seconds_to_target"],
        "ratio_with_chain": best(unis, wc)[wc] / best(lrns, wc)[wc],
        "agreement_sigmas": abs(uni["beta_mu"] - lrn["beta_mu"]) / math.hypot(uni["sigma"], lrn["sigma"]),
        "target_sigma": target,
        "insertion_variance_gain_fixed_configuration": var_u / var_l if var_l > 0 else None,
        "uniform": uni, "learned": lrn,
        "counts_tried": {"uniform": [brief(d) for d in unis], "learned": [brief(d) for d in lrns]},
        "chain": {"gap": gap, "steps": S * gap, "seconds": t_chain},
    }


def _widom_system(rho, melt, tuning=""):
    """N = 108000 (fcc 4*30^3) at density rho, melted by `melt` sweeps."""
    from hsmc_amd import gpu
    conf, box = gpu.lattice(2, 30, rho)
    ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=11)])
    ctx.set_tuning(*((int(v) for v in tuning.split(",")) if tuning else (32, 2, 64)))
    ctx.upload(conf)
    ctx.sweeps(melt)   # the superheated fcc crystal melts (the mean square displacement grows linearly)
    ctx.sync()
    return ctx, box


def _widom_kernel_roofline(res, n_part, grid_cells):
    """Algorithmic bytes of a Widom call (the configuration's positions,
    cell counts and ids read once: N*24 + C*4 + N*4 per sample, amortised
    over its insertions) over the measured time per call, against HBM peak."""
    algo = n_part * 28 + grid_cells * 4
    out = {}
    for name in ("uniform", "learned"):
        ms = res[name]["ms_per_configuration"]
        gbs = algo / (ms / 1e3) / 1e9
        out[name] = {"bound": "hbm", "algorithmic_bytes_per_call": algo, "achieved": gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "insertions_per_s": res[name]["insertions_per_s"]}
    return out


def widom_headline(args) -> dict:
    """The Widom half of the headline metric: GPU seconds to sigma(beta
    mu_ex) = 0.01 at rho = 0.90, N = 108000, learned against uniform, along
    one running NVT chain (4*10^4 melt sweeps, then --widom-configs samples
    --widom-gap sweeps apart)."""
    from hsmc_amd import analysis as A
    ctx, box = _widom_system(0.90, 40000, args.tuning)
    print(f"[bench] widom: N=108000 rho=0.90 melted, box {box[0]:.4f}", file=sys.stderr, flush=True)
    res = _widom_chain(ctx, ctx.sweeps, args.sigma, args.widom_configs, args.widom_gap,
                       40_000_000, 1_000_000)
    ctx.close()
    print(f"[bench] widom: uniform {res['uniform']}", file=sys.stderr, flush=True)
    print(f"[bench] widom: learned {res['learned']}", file=sys.stderr, flush=True)
    g = int(box[0] // 1.0)
    res["workload"] = ("N=108000 (fcc 4*30^3) rho=0.90 melted by 4*10^4 sweeps; NVT chain, one "
                       "sample every gap sweeps; each method at each of its insertion counts per "
                       "sample on the same configurations (uniform, cell-stratified: about 1.6e8, "
                       "4e7, 1e7, 2.5e6 points; learned: 4e6, 1e6, 2.5e5 insertions, each call "
                       "scoring its configuration) and reported at its fastest; voxel 0.25 sigma "
                       "(198^3), mix "
                       "0.9, network 12-64-64-1 trained on 1e6 uniform insertions (training "
                       "charged to the learned seconds)")
    res["carnahan_starling_beta_mu_ex"] = A.carnahan_starling(0.90)["mu_ex"]
    res["roofline"] = _widom_kernel_roofline(res, 108000, g ** 3)
    return res


def _load_state(ctx, path) -> bool:
    """Upload a configuration written by _save_state, if the file exists."""
    import numpy as np
    if not path or not os.path.exists(path):
        return False
    ctx.upload(np.load(path))
    print(f"[bench] configuration loaded from {path}", file=sys.stderr, flush=True)
    return True


def _save_state(ctx, args) -> bool:
    """Write the melted configuration to --state; True when the run should
    stop there (--save-only)."""
    import numpy as np
    if args.state:
        np.save(args.state, ctx.download())
        print(f"[bench] configuration written to {args.state}", file=sys.stderr, flush=True)
    return bool(args.state and args.save_only)


def profile_widom(args) -> None:
    """Only the Widom kernels on the rho = 0.90, N = 108000 system (a few
    calls each), for rocprofv3 counter runs."""
    from hsmc_amd import gpu
    if args.state and os.path.exists(args.state):
        ctx, _ = _widom_system(0.90, 0, args.tuning)
        _load_state(ctx, args.state)
    else:
        ctx, _ = _widom_system(0.90, 2000, args.tuning)
        if _save_state(ctx, args):
            ctx.close()
            return
    for k in range(3):
        ctx.widom_uniform(5, k, 0, 100_000_000)
        ctx.widom_cells(6, k, 850)   # about 1e8 points, cell-stratified
        ctx.widom_cells(7, k, 85)    # about 1e7 points
    net = gpu.Mlp(ctx, seed=5)
    net.train(0.25, 7, 0, 100_000, epochs=1, batch=4096, lr=2e-3, beta=20.0)
    for k in range(3):
        ctx.sweeps(1)   # a new configuration: the learned call scores it
        net.widom(0.25, 0.9, 9, k, 1_000_000)
    ctx.sync()
    net.close()
    ctx.close()
    print("[bench] widom kernels done", file=sys.stderr, flush=True)


def widom(args) -> None:
    """Learned against uniform Widom at rho = 0.70, 0.80, 0.90 (N = 108000,
    voxel 0.25 sigma, mix 0.9), each along its own running NVT chain; the
    headline value is rho = 0.90.  Insertions per sample are sized to each
    density's insertion probability."""
    from hsmc_amd import analysis as A
    sizes = {0.70: (2_000_000, 200_000), 0.80: (20_000_000, 500_000), 0.90: (100_000_000, 1_000_000)}
    res = {}
    for rho in (0.70, 0.80, 0.90):
        ctx, _ = _widom_system(rho, 10000, args.tuning)
        out = _widom_chain(ctx, ctx.sweeps, args.sigma, args.widom_configs, args.widom_gap, *sizes[rho])
        ctx.close()
        out["carnahan_starling_beta_mu_ex"] = A.carnahan_starling(rho)["mu_ex"]
        print(f"[bench] widom rho {rho}: ratio {out['ratio']:.2f}, agreement "
              f"{out['agreement_sigmas']:.2f} sigma", file=sys.stderr, flush=True)
        res[str(rho)] = out
    top = res["0.9"]
    print(json.dumps({
        "metric": "GPU seconds to sigma(beta mu_ex)=0.01 at rho=0.90, learned vs uniform Widom",
        "value": top["ratio"], "unit": "x less GPU time (uniform / learned)",
        "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "f64 (overlap), bf16 (network)",
        "data": "synthetic (fcc 4*30^3 melted by 10^4 sweeps; one running NVT chain per density)",
        "config": {"workload": "N=108000, rho in {0.70, 0.80, 0.90}, voxel 0.25 sigma, mix 0.9, "
                               "network 12-64-64-1 trained on 1e6 uniform insertions",
                   "target_sigma": args.sigma, "configurations": args.widom_configs,
                   "gap_sweeps": args.widom_gap},
        "densities": res,
    }))

def _cs_density(p):
    """Carnahan-Starling density at reduced pressure beta P sigma^3 = p."""
    from hsmc_amd import analysis as A
    lo, hi = 0.1, 0.99
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if mid * A.carnahan_starling(mid)["Z"] < p:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def widom_npt(args) -> None:
    """Learned against uniform Widom on NpT chains of N = 256000 at beta P =
    5, 8, 11.  Per pressure one system starts from fcc 4*40^3 at the
    Carnahan-Starling density of that pressure, melts in NVT sweeps
    (--warmup; the crystal at rho 0.93 needs about 1.6e4 to start
    diffusing), runs NpT sweeps (--steps, ln-volume step 0.001), and then
    the Widom chain continues with NpT sweeps between samples.  At N = 256000
    an ln-volume step of up to 5e-4 compresses hundreds of contact pairs into
    overlap, so about 0.1% of volume moves are accepted and the density stays
    near its start.  The headline value is beta P = 11."""
    import numpy as np

    from hsmc_amd import analysis as A
    from hsmc_amd import gpu
    levels, n_cells, dv = [5.0, 8.0, 11.0], 40, 0.001
    melt, settle = args.warmup, args.steps
    sizes = {5.0: (2_000_000, 200_000), 8.0: (25_000_000, 500_000), 11.0: (200_000_000, 1_000_000)}
    res = {}
    for p in levels:
        conf, box = gpu.lattice(2, n_cells, _cs_density(p))
        ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=500 + int(p))])
        ctx.set_tuning(*((int(v) for v in args.tuning.split(",")) if args.tuning else (32, 2, 64)))
        ctx.upload(conf)
        ctx.sweeps(melt)
        vol = ctx.npt_sweeps(settle, [p], dv)
        advance = lambda g, p=p: ctx.npt_sweeps(g, [p], dv)
        out = _widom_chain(ctx, advance, args.sigma, args.widom_configs, args.widom_gap, *sizes[p],
                           seed=int(p))
        if ctx.any_overlap_scaled([1.0])[0] != 1:
            raise SystemExit("overlap found after the NpT sweeps")
        dens = len(conf) / float(np.prod(ctx.box()))
        ctx.close()
        out.update(beta_P=p, density=dens, carnahan_starling_density=_cs_density(p),
                   carnahan_starling_beta_mu_ex=A.carnahan_starling(dens)["mu_ex"],
                   settle_volume_moves_accepted=f"{int(vol[0, 1])} of {int(vol[0, 0])}")
        print(f"[bench] widom_npt beta P {p}: density {dens:.5f}, ratio {out['ratio']:.2f}, "
              f"agreement {out['agreement_sigmas']:.2f} sigma", file=sys.stderr, flush=True)
        res[str(p)] = out
    top = res["11.0"]
    print(json.dumps({
        "metric": "GPU seconds to sigma(beta mu_ex)=0.01 on NpT chains at beta P = 11, "
                  "learned vs uniform Widom",
        "value": top["ratio"], "unit": "x less GPU time (uniform / learned)",
        "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "f64 (overlap), bf16 (network)",
        "data": f"synthetic (fcc 4*40^3 at the Carnahan-Starling density of each pressure, {melt} NVT "
                f"sweeps, {settle} NpT sweeps, then NpT sweeps between Widom samples)",
        "config": {"workload": "N=256000 NpT chains at beta P in {5, 8, 11}, voxel 0.25 sigma, mix 0.9, "
                               "network 12-64-64-1 trained on 1e6 uniform insertions per pressure",
                   "target_sigma": args.sigma, "melt_sweeps": melt, "npt_sweeps": settle,
                   "configurations": args.widom_configs, "gap_sweeps": args.widom_gap},
        "pressures": res,
    }))


if __name__ == "__main__":
    main()
