"""Write the golden vectors under tests/golden/ from fedluc/HSMC's own code.

Runs oracle/_ref/ref_harness -- the reference's hot-path translation units
compiled from their sources (oracle/ref/Makefile) -- on chosen configurations,
trial positions and scripted moves, and stores inputs and the reference's
outputs in tests/golden/golden.npz.  Only data is written.  This script needs
the reference tree and runs in the build container; the tests read the .npz.

Configurations are inputs, not outputs: the lattices come from the reference
itself; the disordered ones are MC snapshots made by the oracle's serial sweep
(any valid hard-sphere configuration serves).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run(cmd, sysargs, conf, *args, tmp):
    confp = "-"
    if conf is not None:
        confp = os.path.join(tmp, "conf.bin")
        np.ascontiguousarray(conf, dtype=np.float64).tofile(confp)
    argv = [HARNESS, cmd] + [str(a) for a in sysargs] + [confp] + [str(a) for a in args]
    subprocess.run(argv, check=True, stdout=subprocess.DEVNULL)


def system_case(name, type_, n, rho, neigh_dr, max_part, sweeps, seed, out, tmp, rng):
    """Lattice or MC snapshot plus every reference observable on it."""
    sysargs = (type_, n, n, n, repr(rho), repr(neigh_dr), max_part)
    o = O.System(type_, n, rho=rho, neigh_dr=neigh_dr, max_part=max_part, dr_max=0.1, seed=seed)
    if sweeps:
        o.sweep_nvt(sweeps)
    conf = o.config()
    N = conf.shape[0]
    box = o.box()
    L = box["lx"]
    out[f"{name}/sys"] = np.array([type_, n, rho, neigh_dr, max_part], dtype=np.float64)
    out[f"{name}/conf"] = conf

    # box and lattice from the reference (sim_box_init, part_init)
    run("box", sysargs, None, os.path.join(tmp, "b.bin"), os.path.join(tmp, "lat.bin"), tmp=tmp)
    out[f"{name}/ref_box"] = np.fromfile(os.path.join(tmp, "b.bin"))
    if not sweeps:
        out[f"{name}/ref_lattice"] = np.fromfile(os.path.join(tmp, "lat.bin")).reshape(N, 4)

    # fresh cell list (cell_list_new) on the snapshot
    run("celllist", sysargs, conf, os.path.join(tmp, "pc.bin"), os.path.join(tmp, "nm.bin"), tmp=tmp)
    nc = int(np.floor(L / neigh_dr))
    out[f"{name}/ref_part_cell"] = np.fromfile(os.path.join(tmp, "pc.bin"), dtype=np.int32).reshape(nc ** 3, max_part)
    out[f"{name}/ref_neigh_mat"] = np.fromfile(os.path.join(tmp, "nm.bin"), dtype=np.int32).reshape(nc ** 3, 27)

    # overlap decisions for trial positions (check_overlap after a move)
    m = 2000
    idx = rng.integers(0, N, m)
    scale = rng.choice([0.05, 0.1, 0.3, 0.6, 1.2], m)
    trial = conf[idx, 1:] + (rng.random((m, 3)) - 0.5) * scale[:, None]
    trial = np.where(trial > L, trial - L, np.where(trial < 0.0, trial + L, trial))
    q = np.column_stack([idx.astype(np.float64), trial])
    q.tofile(os.path.join(tmp, "q.bin"))
    run("overlap", sysargs, conf, os.path.join(tmp, "q.bin"), os.path.join(tmp, "o.bin"), tmp=tmp)
    out[f"{name}/overlap_q"] = q
    out[f"{name}/overlap_ref"] = np.fromfile(os.path.join(tmp, "o.bin"), dtype=np.uint8)

    # scaled overlap (vol_move / presst building block) and compute_dist
    m = 2000
    sfs = rng.choice([1.0, 0.9999, 0.99966, 0.9993, 1.0002, 0.995], m)
    q = np.column_stack([rng.integers(0, N, m).astype(np.float64), sfs])
    q.tofile(os.path.join(tmp, "q.bin"))
    run("scaled", sysargs, conf, os.path.join(tmp, "q.bin"), os.path.join(tmp, "o.bin"), tmp=tmp)
    out[f"{name}/scaled_q"] = q
    out[f"{name}/scaled_ref"] = np.fromfile(os.path.join(tmp, "o.bin"), dtype=np.uint8)
    q = np.column_stack([rng.integers(0, N, m).astype(np.float64),
                         rng.integers(0, N, m).astype(np.float64), sfs])
    q.tofile(os.path.join(tmp, "q.bin"))
    run("dist", sysargs, conf, os.path.join(tmp, "q.bin"), os.path.join(tmp, "o.bin"), tmp=tmp)
    out[f"{name}/dist_q"] = q
    out[f"{name}/dist_ref"] = np.fromfile(os.path.join(tmp, "o.bin"))

    # rdf (rdf_hist_compute, rdf_hist_norm) at 0.01 up to 5 (clamped to L/2)
    run("rdf", sysargs, conf, "0.01", "5", os.path.join(tmp, "h.bin"), os.path.join(tmp, "g.bin"), tmp=tmp)
    h = np.fromfile(os.path.join(tmp, "h.bin"))
    out[f"{name}/rdf_hist_ref"] = h[:-1]
    out[f"{name}/rdf_rmax_ref"] = h[-1:]
    out[f"{name}/rdf_g_ref"] = np.fromfile(os.path.join(tmp, "g.bin"))

    # contact histogram (pressv_compute_hist) needs cells >= 1.05
    if L / nc >= 1.05:
        run("contact", sysargs, conf, "0.01", os.path.join(tmp, "h.bin"), os.path.join(tmp, "g.bin"), tmp=tmp)
        h = np.fromfile(os.path.join(tmp, "h.bin"))
        out[f"{name}/contact_hist_ref"] = h[:-1]
        out[f"{name}/contact_rmax_ref"] = h[-1:]
        out[f"{name}/contact_g_ref"] = np.fromfile(os.path.join(tmp, "g.bin"))

    # thermodynamic pressure flags (presst_compute_hist), default dxi/xi_max
    # and a coarser one that reaches overlaps
    for tag, dxi, xim in (("a", "0.0001", "0.002"), ("b", "0.002", "0.04")):
        run("presst", sysargs, conf, dxi, xim, os.path.join(tmp, "h.bin"), tmp=tmp)
        out[f"{name}/presst_{tag}_ref"] = np.fromfile(os.path.join(tmp, "h.bin"))
        out[f"{name}/presst_{tag}_args"] = np.array([float(dxi), float(xim)])

    # Widom trial points, some on the faces of the box
    m = 4000
    pts = rng.random((m, 3)) * L
    pts[:50, 0] = 0.0
    pts[50:100, 1] = 0.0
    pts.tofile(os.path.join(tmp, "q.bin"))
    run("widom", sysargs, conf, os.path.join(tmp, "q.bin"), os.path.join(tmp, "o.bin"), tmp=tmp)
    out[f"{name}/widom_q"] = pts
    out[f"{name}/widom_ref"] = np.fromfile(os.path.join(tmp, "o.bin"), dtype=np.uint8)

    # scripted part_move sequence (incremental cell list, history order)
    m = min(20 * N, 16000)
    mt = O.MT(seed + 17)
    moves = np.zeros((m, 4))
    for k in range(m):
        moves[k, 0] = mt.int(N)
        moves[k, 1] = mt.double()
        moves[k, 2] = mt.double()
        moves[k, 3] = mt.double()
    moves.tofile(os.path.join(tmp, "q.bin"))
    run("moves", sysargs, conf, os.path.join(tmp, "q.bin"), "0.1", os.path.join(tmp, "a.bin"),
        os.path.join(tmp, "c.bin"), os.path.join(tmp, "pc.bin"), tmp=tmp)
    out[f"{name}/moves_q"] = moves
    out[f"{name}/moves_acc_ref"] = np.fromfile(os.path.join(tmp, "a.bin"), dtype=np.uint8)
    out[f"{name}/moves_conf_ref"] = np.fromfile(os.path.join(tmp, "c.bin")).reshape(N, 4)
    out[f"{name}/moves_part_cell_ref"] = np.fromfile(os.path.join(tmp, "pc.bin"), dtype=np.int32).reshape(nc ** 3, max_part)


PARSE_CASES = {
    "c1": """# reference-style input, NVT at rho 0.5
rho 0.5
cells_x 5
cells_y 5
cells_z 5
type 2
neigh_list 1.05 10
dr_max 0.1
opt 0 1000 10 0.5 0.5
seed 1
sweep_eq 2000
sweep_stat 20000
out 1000
press_virial 0.01 10
rdf 0.01 5 10 100
widom 100 10
""",
    "npt": """rho 0.75
cells_x 6
cells_y 6
cells_z 6
type 2
npt 5.0 0.001
press_thermo 0.0001 0.002 5
config_write 100 10
restart_write 1000
seed 42
sweep_eq 10
sweep_stat 20
out 5
""",
    "defaults": """# only comments and blank lines

""",
}


def main() -> None:
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    rng = np.random.default_rng(20261202)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        system_case("c1_lattice", 2, 5, 0.5, 1.05, 10, 0, 1, out, tmp, rng)
        system_case("c1_mc", 2, 5, 0.5, 1.05, 10, 200, 3, out, tmp, rng)
        system_case("c1_mc_dr1", 2, 5, 0.5, 1.0, 10, 150, 5, out, tmp, rng)
        system_case("dense_sc", 1, 9, 0.85, 1.0, 10, 100, 7, out, tmp, rng)
        system_case("n4000", 2, 10, 0.85, 1.05, 10, 60, 11, out, tmp, rng)
        for name, text in PARSE_CASES.items():
            path = os.path.join(HERE, f"parse_{name}.in")
            with open(path, "w") as f:
                f.write(text)
            run("parse", (2, 3, 3, 3, "0.5", "1.0", 10), None, path, os.path.join(tmp, "p.bin"), tmp=tmp)
            out[f"parse/{name}"] = np.fromfile(os.path.join(tmp, "p.bin"))
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print(f"wrote {len(out)} arrays to tests/golden/golden.npz")


if __name__ == "__main__":
    main()
