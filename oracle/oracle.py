"""ctypes front end of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, always as the checker or as the timed
CPU restatement of the reference, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile liborc.so with the Makefile beside this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, u64, i64, f64 = C.c_void_p, C.c_int, C.c_uint64, C.c_int64, C.c_double
        P = C.POINTER
        sig = {
            "orc_mt_new": (vp, [C.c_ulong]),
            "orc_mt_free": (None, [vp]),
            "orc_mt_next": (C.c_uint32, [vp]),
            "orc_mt_double": (f64, [vp]),
            "orc_mt_int": (i32, [vp, i32]),
            "orc_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
            "orc_create": (vp, [i32, i32, i32, i32, f64, f64, i32, f64, C.c_ulong, P(i32)]),
            "orc_destroy": (None, [vp]),
            "orc_set_params": (None, [vp, f64, f64, f64]),
            "orc_set_seed": (None, [vp, C.c_ulong]),
            "orc_set_config": (i32, [vp, P(f64)]),
            "orc_get_config": (None, [vp, P(f64)]),
            "orc_get_cell_list": (None, [vp, P(i32), P(i32)]),
            "orc_get_box": (None, [vp, P(f64)]),
            "orc_get_counters": (None, [vp, P(u64)]),
            "orc_reset_counters": (None, [vp]),
            "orc_cell_list_new": (i32, [vp]),
            "orc_compute_dist": (f64, [vp, i32, i32, f64, f64, f64]),
            "orc_check_overlap": (i32, [vp, i32, f64, f64, f64]),
            "orc_overlap_at": (i32, [vp, i32, f64, f64, f64]),
            "orc_part_move": (i32, [vp]),
            "orc_part_move_scripted": (i32, [vp, i32, f64, f64, f64]),
            "orc_vol_move": (i32, [vp]),
            "orc_sweep_nvt": (i32, [vp, i32]),
            "orc_sweep_npt": (i32, [vp, i32]),
            "orc_rdf_counts": (i32, [vp, f64, i32, f64, P(i64), P(i64)]),
            "orc_contact_counts": (i32, [vp, f64, i32, f64, P(i64), P(i64)]),
            "orc_presst_flags": (i32, [vp, P(f64), i32, P(C.c_uint8)]),
            "orc_widom_points": (i64, [vp, P(f64), i32, P(C.c_uint8)]),
            "orc_widom_insertion": (i64, [vp, i32]),
            "orc_widom_philox": (i64, [vp, u64, u64, i64, i64]),
            "orc_widom_cells": (i64, [vp, u64, u64, i32, i32, i32, i32]),
            "orc_widom_philox_point": (None, [vp, u64, u64, i64, P(f64)]),
            "orc_time_sweeps": (i32, [vp, i32, P(f64), P(f64)]),
            "orc_cb_create": (vp, [i32, P(f64), f64, f64, f64, i32, i32, i32, f64, u64]),
            "orc_cb_destroy": (None, [vp]),
            "orc_cb_sweeps": (i32, [vp, i32]),
            "orc_cb_get_config": (None, [vp, P(f64)]),
            "orc_cb_get_counters": (None, [vp, P(u64)]),
            "orc_cb_sweep_randoms": (None, [vp, u64, P(f64), P(i32)]),
            "orc_cb_pick": (C.c_uint32, [P(C.c_uint32), C.c_uint32, C.c_uint32, u64, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class MT:
    """MT19937 with the reference's GSL conventions (rng.c:21-36)."""

    def __init__(self, seed: int):
        self.h = lib().orc_mt_new(seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mt_free(self.h)
            self.h = None

    def next(self) -> int:
        return lib().orc_mt_next(self.h)

    def double(self) -> float:
        return lib().orc_mt_double(self.h)

    def int(self, n: int) -> int:
        return lib().orc_mt_int(self.h, n)


def philox(ctr, key) -> tuple:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return tuple(int(v) for v in out)


def cb_pick(seed: int, cell: int, t: int, sweep: int, n: int) -> int:
    """The checkerboard sweep's occupant pick (orc_cb_pick)."""
    k = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], dtype=np.uint32)
    return int(lib().orc_cb_pick(_p(k, C.c_uint32), cell, t, sweep, n))


class OracleError(RuntimeError):
    pass


class System:
    """One reference-style system: box, lattice, cell list, MT stream."""

    def __init__(self, type: int = 2, nx: int = 5, ny: int | None = None, nz: int | None = None,
                 rho: float = 0.5, neigh_dr: float = 1.0, max_part: int = 10,
                 dr_max: float = 0.1, seed: int = 1):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        st = C.c_int(0)
        self.h = lib().orc_create(type, nx, ny, nz, rho, neigh_dr, max_part, dr_max, seed,
                                  C.byref(st))
        if not self.h:
            raise OracleError(f"orc_create failed with status {st.value}")
        self.type, self.nx, self.ny, self.nz = type, nx, ny, nz
        self.max_part = max_part
        self.neigh_dr = neigh_dr
        self.N = nx * ny * nz * (1 if type == 1 else 4)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    # --- state ---
    def box(self) -> dict:
        b = np.zeros(8)
        lib().orc_get_box(self.h, _p(b, C.c_double))
        return dict(vol=b[0], lx=b[1], ly=b[2], lz=b[3], size_x=b[4], size_y=b[5],
                    size_z=b[6], rho=b[7])

    def config(self) -> np.ndarray:
        out = np.zeros((self.N, 4))
        lib().orc_get_config(self.h, _p(out, C.c_double))
        return out

    def set_config(self, conf: np.ndarray) -> None:
        c = np.ascontiguousarray(conf, dtype=np.float64)
        assert c.shape == (self.N, 4)
        st = lib().orc_set_config(self.h, _p(c, C.c_double))
        if st:
            raise OracleError(f"orc_set_config status {st}")

    def ncells(self) -> tuple:
        b = self.box()
        import math
        n = int(math.floor(b["lx"] / self.neigh_dr))
        return n

    def cell_list(self) -> tuple:
        n = self.ncells()
        C_ = n ** 3
        pc = np.zeros((C_, self.max_part), dtype=np.int32)
        nm = np.zeros((C_, 27), dtype=np.int32)
        lib().orc_get_cell_list(self.h, _p(pc, C.c_int), _p(nm, C.c_int))
        return pc, nm

    def set_params(self, dr_max: float, dv_max: float = 0.001, press: float = 0.0) -> None:
        lib().orc_set_params(self.h, dr_max, dv_max, press)

    def set_seed(self, seed: int) -> None:
        lib().orc_set_seed(self.h, seed)

    def counters(self) -> np.ndarray:
        out = np.zeros(6, dtype=np.uint64)
        lib().orc_get_counters(self.h, _p(out, C.c_uint64))
        return out

    def reset_counters(self) -> None:
        lib().orc_reset_counters(self.h)

    # --- moves ---
    def compute_dist(self, i: int, j: int, sf: float = 1.0) -> float:
        return lib().orc_compute_dist(self.h, i, j, sf, sf, sf)

    def check_overlap(self, i: int, sf: float = 1.0) -> int:
        return lib().orc_check_overlap(self.h, i, sf, sf, sf)

    def overlap_at(self, i: int, x: float, y: float, z: float) -> int:
        return lib().orc_overlap_at(self.h, i, x, y, z)

    def part_move_scripted(self, idx: int, rx: float, ry: float, rz: float) -> int:
        return lib().orc_part_move_scripted(self.h, idx, rx, ry, rz)

    def sweep_nvt(self, n: int = 1) -> None:
        st = lib().orc_sweep_nvt(self.h, n)
        if st:
            raise OracleError(f"sweep_nvt status {st}")

    def sweep_npt(self, n: int = 1) -> None:
        st = lib().orc_sweep_npt(self.h, n)
        if st:
            raise OracleError(f"sweep_npt status {st}")

    def vol_move(self) -> int:
        return lib().orc_vol_move(self.h)

    def time_sweeps(self, n: int) -> tuple:
        w = C.c_double(0)
        c = C.c_double(0)
        st = lib().orc_time_sweeps(self.h, n, C.byref(w), C.byref(c))
        if st:
            raise OracleError(f"time_sweeps status {st}")
        return w.value, c.value

    # --- observables ---
    def rdf_counts(self, dr_bin: float, nbins: int, rmax_eff: float) -> tuple:
        out = np.zeros(nbins, dtype=np.int64)
        oor = C.c_int64(0)
        lib().orc_rdf_counts(self.h, dr_bin, nbins, rmax_eff, _p(out, C.c_int64), C.byref(oor))
        return out, oor.value

    def contact_counts(self, dr_bin: float, nbins: int, rmax_eff: float) -> tuple:
        out = np.zeros(nbins, dtype=np.int64)
        oor = C.c_int64(0)
        lib().orc_contact_counts(self.h, dr_bin, nbins, rmax_eff, _p(out, C.c_int64),
                                 C.byref(oor))
        return out, oor.value

    def presst_flags(self, sf) -> np.ndarray:
        s = np.ascontiguousarray(sf, dtype=np.float64)
        out = np.zeros(len(s), dtype=np.uint8)
        lib().orc_presst_flags(self.h, _p(s, C.c_double), len(s), _p(out, C.c_uint8))
        return out

    def widom_points(self, pts: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(p), dtype=np.uint8)
        lib().orc_widom_points(self.h, _p(p, C.c_double), len(p), _p(out, C.c_uint8))
        return out

    def widom_insertion(self, n: int) -> int:
        return lib().orc_widom_insertion(self.h, n)

    def widom_philox(self, seed: int, sample: int, first: int, n: int) -> int:
        return lib().orc_widom_philox(self.h, seed, sample, first, n)

    def widom_cells(self, seed: int, sample: int, grid, per_cell: int) -> int:
        """Hits of the cell-stratified insertion (restates hsmc_gpu_widom_cells)."""
        nx, ny, nz = grid
        return int(lib().orc_widom_cells(self.h, seed, sample, nx, ny, nz, per_cell))

    def widom_philox_point(self, seed: int, sample: int, i: int) -> np.ndarray:
        out = np.zeros(3)
        lib().orc_widom_philox_point(self.h, seed, sample, i, _p(out, C.c_double))
        return out


class Checkerboard:
    """CPU restatement of the HIP checkerboard sweep (same Philox stream)."""

    def __init__(self, conf: np.ndarray, box, grid, dr_max: float, seed: int):
        c = np.ascontiguousarray(conf, dtype=np.float64)
        self.N = c.shape[0]
        lx, ly, lz = box
        gx, gy, gz = grid
        self.h = lib().orc_cb_create(self.N, _p(c, C.c_double), lx, ly, lz, gx, gy, gz,
                                     dr_max, seed)
        if not self.h:
            raise OracleError("orc_cb_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cb_destroy(self.h)
            self.h = None

    def sweeps(self, n: int) -> None:
        st = lib().orc_cb_sweeps(self.h, n)
        if st:
            raise OracleError(f"cb sweeps status {st}")

    def config(self) -> np.ndarray:
        out = np.zeros((self.N, 4))
        lib().orc_cb_get_config(self.h, _p(out, C.c_double))
        return out

    def counters(self) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint64)
        lib().orc_cb_get_counters(self.h, _p(out, C.c_uint64))
        return out

    def sweep_randoms(self, sweep: int) -> tuple:
        sh = np.zeros(3)
        perm = np.zeros(8, dtype=np.int32)
        lib().orc_cb_sweep_randoms(self.h, sweep, _p(sh, C.c_double), _p(perm, C.c_int))
        return sh, perm
