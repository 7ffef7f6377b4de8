"""Batched replicas: many independent systems of uneven size and density in
one context sweep together; each must evolve exactly as it would alone."""
import numpy as np
import pytest

from hsmc_amd import gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _replicas(count, seed, nmin=3, nmax=8):
    rng = np.random.default_rng(seed)
    reps = []
    for r in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        rho = float(rng.uniform(0.30, 0.93))
        conf, box = gpu.lattice(2, n, rho)
        reps.append(dict(n=len(conf), box=box, dr_max=0.12, seed=7000 + 13 * r, conf=conf, rho=rho))
    return reps


def test_batched_replicas_each_match_oracle():
    reps = _replicas(16, 5)
    ctx = gpu.Context(reps)
    for r, rep in enumerate(reps):
        ctx.upload(rep["conf"], r)
    ctx.sweeps(5)
    for r, rep in enumerate(reps):
        cb = O.Checkerboard(rep["conf"], rep["box"], ctx.grid(r), 0.12, rep["seed"])
        cb.sweeps(5)
        assert np.array_equal(ctx.download(r), cb.config()), r
        c = ctx.counters(r)
        ref = cb.counters()
        assert (c["attempts"], c["accepted"], c["left_cell"]) == tuple(int(v) for v in ref[:3]), r


def test_batched_equals_alone():
    reps = _replicas(6, 9)
    together = gpu.Context(reps)
    for r, rep in enumerate(reps):
        together.upload(rep["conf"], r)
    together.set_tuning(64, 2, 4)
    together.sweeps(9)
    for r, rep in enumerate(reps):
        alone = gpu.Context([rep])
        alone.upload(rep["conf"])
        alone.sweeps(9)
        assert np.array_equal(together.download(r), alone.download())


def test_batched_observables_per_replica():
    reps = _replicas(4, 11, 5, 7)
    ctx = gpu.Context(reps)
    for r, rep in enumerate(reps):
        ctx.upload(rep["conf"], r)
    ctx.sweeps(20)
    for r, rep in enumerate(reps):
        conf = ctx.download(r)
        n = round((rep["n"] / 4) ** (1 / 3))
        o = O.System(2, n, rho=rep["rho"], neigh_dr=1.05, max_part=16)
        o.set_config(conf)
        L = rep["box"][0]
        nn = int((min(5.0, L / 2) - 1.0) / 0.01)
        rmax = 0.01 * nn + 1.0
        assert np.array_equal(ctx.rdf_counts(0.01, nn, rmax, r)[0], o.rdf_counts(0.01, nn, rmax)[0])
        assert ctx.any_overlap_scaled([1.0], r)[0] == 1


def test_autotune_keeps_state():
    reps = _replicas(8, 3)
    ctx = gpu.Context(reps)
    for r, rep in enumerate(reps):
        ctx.upload(rep["conf"], r)
    ctx.sweeps(3)
    before = [ctx.download(r) for r in range(len(reps))]
    c0 = ctx.counters(0)
    best = ctx.autotune(8)
    assert best["lanes_per_cell"] in (32, 64)
    after = [ctx.download(r) for r in range(len(reps))]
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
    assert ctx.counters(0) == c0
