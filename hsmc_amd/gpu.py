"""ctypes binding of libhsmc_gpu.so (the C ABI declared in include/hsmc_gpu.h).

This is a thin host-side mirror for Python callers (tests, bench, analysis
drivers).  There is no CPU fallback: every call goes to the HIP library, and
loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhsmc_gpu.so")

HSMC_OK = 0
ERRORS = {
    -1: "CELL_OVERFLOW", -2: "CELL_TOO_SMALL", -3: "INIT_OVERLAP", -4: "HIP",
    -5: "BAD_ARG", -7: "NO_MEMORY", -8: "NOT_SUPPORTED", -9: "NO_DEVICE",
}

EXPORTS = [
    "hsmc_gpu_create", "hsmc_gpu_destroy", "hsmc_gpu_strerror", "hsmc_gpu_last_error",
    "hsmc_gpu_sync", "hsmc_gpu_upload", "hsmc_gpu_download", "hsmc_gpu_set_box",
    "hsmc_gpu_scale", "hsmc_gpu_grid", "hsmc_gpu_sweeps", "hsmc_gpu_time_sweeps",
    "hsmc_gpu_counters", "hsmc_gpu_reset_counters", "hsmc_gpu_set_sweep_counter",
    "hsmc_gpu_set_tuning", "hsmc_gpu_box", "hsmc_gpu_npt_sweeps",
    "hsmc_gpu_any_overlap_scaled_batch",
    "hsmc_gpu_get_tuning", "hsmc_gpu_autotune", "hsmc_gpu_time_phase_kernel",
    "hsmc_gpu_cell_list", "hsmc_gpu_overlap_batch", "hsmc_gpu_overlap_scaled",
    "hsmc_gpu_any_overlap_scaled", "hsmc_gpu_pair_dist", "hsmc_gpu_rdf_counts",
    "hsmc_gpu_contact_counts", "hsmc_gpu_widom_points", "hsmc_gpu_widom_uniform",
    "hsmc_gpu_set_particles", "hsmc_gpu_set_dr_max", "hsmc_gpu_occupant_picks",
    "hsmc_gpu_any_overlap_scaled_noalloc", "hsmc_gpu_widom_cells",
    "hsmc_gpu_mlp_create", "hsmc_gpu_mlp_destroy", "hsmc_gpu_mlp_set", "hsmc_gpu_mlp_get",
    "hsmc_gpu_mlp_forward", "hsmc_gpu_mlp_gradient", "hsmc_gpu_point_descriptors",
    "hsmc_gpu_mlp_train", "hsmc_gpu_voxel_scores", "hsmc_gpu_proposal_scores", "hsmc_gpu_mlp_forget",
    "hsmc_gpu_widom_learned",
]

MLP_IN, MLP_HIDDEN = 12, 64
MLP_NPARAM = 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1


class Replica(C.Structure):
    _fields_ = [("n_part", C.c_int32), ("reserved", C.c_int32), ("lx", C.c_double),
                ("ly", C.c_double), ("lz", C.c_double), ("dr_max", C.c_double),
                ("seed", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("lanes_per_cell", C.c_int32), ("cells_per_block", C.c_int32),
                ("sweeps_per_graph", C.c_int32), ("persistent", C.c_int32)]

    def as_dict(self) -> dict:
        return dict(lanes_per_cell=self.lanes_per_cell, cells_per_block=self.cells_per_block,
                    sweeps_per_graph=self.sweeps_per_graph, persistent=self.persistent)


class HsmcError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} {detail}".strip())


_lib = None
# contexts not yet closed: released, networks first, by an exit hook that runs
# before the interpreter tears down module state, so no device object is
# destroyed during interpreter finalization
_live = weakref.WeakSet()


def _close_live() -> None:
    for ctx in list(_live):
        ctx.close()


atexit.register(_close_live)


def load() -> C.CDLL:
    """Load libhsmc_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C hsmc_amd/csrc); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
    P = C.POINTER
    sig = {
        "hsmc_gpu_create": (i32, [P(vp), i32, i32, P(Replica)]),
        "hsmc_gpu_destroy": (i32, [vp]),
        "hsmc_gpu_strerror": (C.c_char_p, [i32]),
        "hsmc_gpu_last_error": (C.c_char_p, [vp]),
        "hsmc_gpu_sync": (i32, [vp]),
        "hsmc_gpu_upload": (i32, [vp, i32, P(f64)]),
        "hsmc_gpu_download": (i32, [vp, i32, P(f64)]),
        "hsmc_gpu_set_box": (i32, [vp, i32, f64, f64, f64]),
        "hsmc_gpu_set_particles": (i32, [vp, i32, P(i32), P(f64), i32]),
        "hsmc_gpu_scale": (i32, [vp, i32, f64, f64, f64, f64]),
        "hsmc_gpu_set_dr_max": (i32, [vp, i32, f64]),
        "hsmc_gpu_grid": (i32, [vp, i32, P(i32)]),
        "hsmc_gpu_sweeps": (i32, [vp, i32]),
        "hsmc_gpu_time_sweeps": (i32, [vp, i32, P(C.c_float)]),
        "hsmc_gpu_counters": (i32, [vp, i32, P(u64)]),
        "hsmc_gpu_reset_counters": (i32, [vp]),
        "hsmc_gpu_set_sweep_counter": (i32, [vp, u64]),
        "hsmc_gpu_box": (i32, [vp, i32, P(f64)]),
        "hsmc_gpu_npt_sweeps": (i32, [vp, i32, P(f64), P(f64), P(u64)]),
        "hsmc_gpu_any_overlap_scaled_batch": (i32, [vp, P(f64), P(C.c_uint8)]),
        "hsmc_gpu_set_tuning": (i32, [vp, P(Tuning)]),
        "hsmc_gpu_get_tuning": (i32, [vp, P(Tuning)]),
        "hsmc_gpu_autotune": (i32, [vp, i32, P(Tuning)]),
        "hsmc_gpu_time_phase_kernel": (i32, [vp, i32, P(C.c_float), P(i32)]),
        "hsmc_gpu_cell_list": (i32, [vp, i32, i32, i32, i32, i32, P(i32), P(i32)]),
        "hsmc_gpu_overlap_batch": (i32, [vp, i32, P(i32), P(f64), i32, P(C.c_uint8)]),
        "hsmc_gpu_overlap_scaled": (i32, [vp, i32, P(i32), P(f64), i32, P(C.c_uint8)]),
        "hsmc_gpu_any_overlap_scaled": (i32, [vp, i32, P(f64), i32, P(C.c_uint8)]),
        "hsmc_gpu_any_overlap_scaled_noalloc": (i32, [vp, i32, P(f64), i32, P(C.c_uint8)]),
        "hsmc_gpu_pair_dist": (i32, [vp, i32, P(i32), P(i32), i32, f64, P(f64)]),
        "hsmc_gpu_rdf_counts": (i32, [vp, i32, f64, i32, f64, P(i64), P(i64)]),
        "hsmc_gpu_contact_counts": (i32, [vp, i32, f64, i32, f64, P(i64), P(i64)]),
        "hsmc_gpu_widom_points": (i32, [vp, i32, P(f64), i32, P(C.c_uint8), P(i64)]),
        "hsmc_gpu_widom_uniform": (i32, [vp, i32, u64, u64, i64, i64, P(i64)]),
        "hsmc_gpu_widom_cells": (i32, [vp, i32, u64, u64, i32, P(i64), P(i64), P(i32)]),
        "hsmc_gpu_mlp_create": (i32, [vp, P(vp), u64]),
        "hsmc_gpu_mlp_destroy": (i32, [vp]),
        "hsmc_gpu_mlp_set": (i32, [vp, P(C.c_float)]),
        "hsmc_gpu_mlp_get": (i32, [vp, P(C.c_float)]),
        "hsmc_gpu_mlp_forward": (i32, [vp, P(C.c_float), i32, P(C.c_float)]),
        "hsmc_gpu_mlp_gradient": (i32, [vp, P(C.c_float), P(C.c_float), i32, P(C.c_float),
                                        P(C.c_float)]),
        "hsmc_gpu_point_descriptors": (i32, [vp, i32, P(f64), i32, P(C.c_float)]),
        "hsmc_gpu_mlp_train": (i32, [vp, i32, f64, u64, u64, i64, i32, i32, C.c_float, C.c_float,
                                     P(C.c_float)]),
        "hsmc_gpu_voxel_scores": (i32, [vp, i32, f64, P(C.c_float), i64, P(i64)]),
        "hsmc_gpu_proposal_scores": (i32, [vp, i32, f64, P(C.c_float), i64, P(i64), P(i64)]),
        "hsmc_gpu_mlp_forget": (i32, [vp]),
        "hsmc_gpu_widom_learned": (i32, [vp, i32, f64, f64, u64, u64, i64, P(f64), P(i64),
                                         P(i64)]),
        "hsmc_gpu_occupant_picks": (i32, [i32, u64, C.c_uint32, u64, C.c_uint32, i32,
                                          P(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Context:
    """One device context holding one or more independent systems (replicas)."""

    def __init__(self, replicas, device: int = 0):
        L = load()
        reps = (Replica * len(replicas))()
        for k, r in enumerate(replicas):
            reps[k].n_part = int(r["n"])
            reps[k].lx, reps[k].ly, reps[k].lz = (float(v) for v in r["box"])
            reps[k].dr_max = float(r.get("dr_max", 0.1))
            reps[k].seed = int(r.get("seed", 1))
        h = C.c_void_p()
        st = L.hsmc_gpu_create(C.byref(h), device, len(replicas), reps)
        if st != HSMC_OK:
            raise HsmcError(st, "hsmc_gpu_create")
        self.h = h
        self.L = L
        self.replicas = [dict(r) for r in replicas]
        self.R = len(replicas)
        self._nets = weakref.WeakSet()   # networks to destroy before the context
        _live.add(self)

    def close(self) -> None:
        if getattr(self, "h", None):
            for m in list(getattr(self, "_nets", ())):
                m.close()
            self.L.hsmc_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ck(self, st: int, where: str) -> None:
        if st != HSMC_OK:
            msg = self.L.hsmc_gpu_last_error(self.h)
            raise HsmcError(st, where, msg.decode() if msg else "")

    def n(self, r: int = 0) -> int:
        return int(self.replicas[r]["n"])

    # --- configuration ---
    def upload(self, conf4: np.ndarray, r: int = 0) -> None:
        c = np.ascontiguousarray(conf4, dtype=np.float64)
        assert c.shape == (self.n(r), 4)
        self._ck(self.L.hsmc_gpu_upload(self.h, r, _p(c, C.c_double)), "upload")

    def download(self, r: int = 0) -> np.ndarray:
        out = np.zeros((self.n(r), 4))
        self._ck(self.L.hsmc_gpu_download(self.h, r, _p(out, C.c_double)), "download")
        return out

    def set_particles(self, idx, xyz, r: int = 0) -> None:
        i = np.ascontiguousarray(idx, dtype=np.int32)
        q = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self._ck(self.L.hsmc_gpu_set_particles(self.h, r, _p(i, C.c_int32), _p(q, C.c_double),
                                               len(i)), "set_particles")

    def set_box(self, box, r: int = 0) -> None:
        self._ck(self.L.hsmc_gpu_set_box(self.h, r, *map(float, box)), "set_box")
        self.replicas[r]["box"] = tuple(box)

    def set_dr_max(self, dr_max: float, r: int = 0) -> None:
        self._ck(self.L.hsmc_gpu_set_dr_max(self.h, r, float(dr_max)), "set_dr_max")
        self.replicas[r]["dr_max"] = float(dr_max)

    def scale(self, sf: float, box, r: int = 0) -> None:
        self._ck(self.L.hsmc_gpu_scale(self.h, r, float(sf), *map(float, box)), "scale")
        self.replicas[r]["box"] = tuple(box)

    def grid(self, r: int = 0) -> tuple:
        g = np.zeros(3, dtype=np.int32)
        self._ck(self.L.hsmc_gpu_grid(self.h, r, _p(g, C.c_int32)), "grid")
        return tuple(int(v) for v in g)

    def box(self, r: int = 0) -> tuple:
        b = np.zeros(3, dtype=np.float64)
        self._ck(self.L.hsmc_gpu_box(self.h, r, _p(b, C.c_double)), "box")
        return tuple(float(v) for v in b)

    # --- sweeps ---
    def sweeps(self, n: int) -> None:
        self._ck(self.L.hsmc_gpu_sweeps(self.h, int(n)), "sweeps")

    def npt_sweeps(self, n: int, press, dv_max) -> np.ndarray:
        """n NpT sweeps of every replica; press and dv_max per replica (or one
        value for all).  Returns per replica (proposed, accepted, rejected for
        overlap, rejected by the rule) volume moves of these sweeps."""
        R = self.R
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(press, dtype=np.float64), (R,)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dv_max, dtype=np.float64), (R,)))
        out = np.zeros((R, 4), dtype=np.uint64)
        self._ck(self.L.hsmc_gpu_npt_sweeps(self.h, int(n), _p(p, C.c_double), _p(d, C.c_double),
                                            _p(out, C.c_uint64)), "npt_sweeps")
        return out

    def time_sweeps(self, n: int) -> float:
        ms = C.c_float(0)
        self._ck(self.L.hsmc_gpu_time_sweeps(self.h, int(n), C.byref(ms)), "time_sweeps")
        return ms.value

    def time_phase_kernel(self, n_sweeps: int) -> tuple:
        ms = C.c_float(0)
        cnt = C.c_int32(0)
        self._ck(self.L.hsmc_gpu_time_phase_kernel(self.h, int(n_sweeps), C.byref(ms),
                                                   C.byref(cnt)), "time_phase_kernel")
        return ms.value, cnt.value

    def sync(self) -> None:
        self._ck(self.L.hsmc_gpu_sync(self.h), "sync")

    def counters(self, r: int = 0) -> dict:
        out = np.zeros(4, dtype=np.uint64)
        self._ck(self.L.hsmc_gpu_counters(self.h, r, _p(out, C.c_uint64)), "counters")
        return dict(attempts=int(out[0]), accepted=int(out[1]), left_cell=int(out[2]),
                    sweeps=int(out[3]))

    def reset_counters(self) -> None:
        self._ck(self.L.hsmc_gpu_reset_counters(self.h), "reset_counters")

    def set_sweep_counter(self, sweep: int) -> None:
        self._ck(self.L.hsmc_gpu_set_sweep_counter(self.h, sweep), "set_sweep_counter")

    def set_tuning(self, lanes_per_cell: int, cells_per_block: int, sweeps_per_graph: int,
                   persistent: int = 0) -> None:
        t = Tuning(lanes_per_cell, cells_per_block, sweeps_per_graph, persistent)
        self._ck(self.L.hsmc_gpu_set_tuning(self.h, C.byref(t)), "set_tuning")

    def get_tuning(self) -> dict:
        t = Tuning()
        self._ck(self.L.hsmc_gpu_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def autotune(self, sweeps_per_trial: int = 64) -> dict:
        t = Tuning()
        self._ck(self.L.hsmc_gpu_autotune(self.h, int(sweeps_per_trial), C.byref(t)), "autotune")
        return t.as_dict()

    # --- cell list and overlap tests ---
    def cell_list(self, nxyz, max_part: int, r: int = 0) -> tuple:
        nx, ny, nz = nxyz
        C_ = nx * ny * nz
        pc = np.zeros((C_, max_part), dtype=np.int32)
        nm = np.zeros((C_, 27), dtype=np.int32)
        self._ck(self.L.hsmc_gpu_cell_list(self.h, r, nx, ny, nz, max_part, _p(pc, C.c_int32),
                                           _p(nm, C.c_int32)), "cell_list")
        return pc, nm

    def overlap_batch(self, idx, xyz, r: int = 0) -> np.ndarray:
        i = np.ascontiguousarray(idx, dtype=np.int32)
        q = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(i), dtype=np.uint8)
        self._ck(self.L.hsmc_gpu_overlap_batch(self.h, r, _p(i, C.c_int32), _p(q, C.c_double),
                                               len(i), _p(out, C.c_uint8)), "overlap_batch")
        return out

    def overlap_scaled(self, idx, sf, r: int = 0) -> np.ndarray:
        i = np.ascontiguousarray(idx, dtype=np.int32)
        s = np.ascontiguousarray(sf, dtype=np.float64)
        out = np.zeros(len(i), dtype=np.uint8)
        self._ck(self.L.hsmc_gpu_overlap_scaled(self.h, r, _p(i, C.c_int32), _p(s, C.c_double),
                                                len(i), _p(out, C.c_uint8)), "overlap_scaled")
        return out

    def any_overlap_scaled(self, sf, r: int = 0) -> np.ndarray:
        s = np.ascontiguousarray(sf, dtype=np.float64)
        out = np.zeros(len(s), dtype=np.uint8)
        self._ck(self.L.hsmc_gpu_any_overlap_scaled(self.h, r, _p(s, C.c_double), len(s),
                                                    _p(out, C.c_uint8)), "any_overlap_scaled")
        return out

    def any_overlap_scaled_batch(self, sf) -> np.ndarray:
        """Per replica r: 1 when no pair overlaps after scaling by sf[r]."""
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(sf, dtype=np.float64), (self.R,)))
        out = np.zeros(self.R, dtype=np.uint8)
        self._ck(self.L.hsmc_gpu_any_overlap_scaled_batch(self.h, _p(s, C.c_double),
                                                          _p(out, C.c_uint8)),
                 "any_overlap_scaled_batch")
        return out

    def pair_dist(self, i, j, sf: float = 1.0, r: int = 0) -> np.ndarray:
        a = np.ascontiguousarray(i, dtype=np.int32)
        b = np.ascontiguousarray(j, dtype=np.int32)
        out = np.zeros(len(a))
        self._ck(self.L.hsmc_gpu_pair_dist(self.h, r, _p(a, C.c_int32), _p(b, C.c_int32), len(a),
                                           float(sf), _p(out, C.c_double)), "pair_dist")
        return out

    # --- observables ---
    def rdf_counts(self, dr_bin: float, nbins: int, rmax_eff: float, r: int = 0) -> tuple:
        out = np.zeros(nbins, dtype=np.int64)
        oor = C.c_int64(0)
        self._ck(self.L.hsmc_gpu_rdf_counts(self.h, r, dr_bin, nbins, rmax_eff,
                                            _p(out, C.c_int64), C.byref(oor)), "rdf_counts")
        return out, oor.value

    def contact_counts(self, dr_bin: float, nbins: int, rmax_eff: float, r: int = 0) -> tuple:
        out = np.zeros(nbins, dtype=np.int64)
        oor = C.c_int64(0)
        self._ck(self.L.hsmc_gpu_contact_counts(self.h, r, dr_bin, nbins, rmax_eff,
                                                _p(out, C.c_int64), C.byref(oor)), "contact_counts")
        return out, oor.value

    def widom_points(self, xyz, r: int = 0) -> np.ndarray:
        q = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(q), dtype=np.uint8)
        hits = C.c_int64(0)
        self._ck(self.L.hsmc_gpu_widom_points(self.h, r, _p(q, C.c_double), len(q),
                                              _p(out, C.c_uint8), C.byref(hits)), "widom_points")
        return out

    def widom_uniform(self, seed: int, sample: int, first: int, n: int, r: int = 0) -> int:
        hits = C.c_int64(0)
        self._ck(self.L.hsmc_gpu_widom_uniform(self.h, r, int(seed), int(sample), int(first),
                                               int(n), C.byref(hits)), "widom_uniform")
        return hits.value

    def widom_cells(self, seed: int, sample: int, per_cell: int, r: int = 0) -> tuple:
        """Cell-stratified uniform insertion: (hits, points, grid shape)."""
        hits = C.c_int64(0)
        npts = C.c_int64(0)
        grid = np.zeros(3, dtype=np.int32)
        self._ck(self.L.hsmc_gpu_widom_cells(self.h, r, int(seed), int(sample), int(per_cell),
                                             C.byref(hits), C.byref(npts), _p(grid, C.c_int32)),
                 "widom_cells")
        return hits.value, npts.value, tuple(int(v) for v in grid)


class Mlp:
    """The learned-proposal network (hsmc_gpu_mlp_*) of one context."""

    def __init__(self, ctx: "Context", seed: int = 1):
        self.ctx = ctx
        self.L = ctx.L
        h = C.c_void_p()
        ctx._ck(self.L.hsmc_gpu_mlp_create(ctx.h, C.byref(h), int(seed)), "mlp_create")
        self.h = h
        ctx._nets.add(self)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.hsmc_gpu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape == (MLP_NPARAM,)
        self.ctx._ck(self.L.hsmc_gpu_mlp_set(self.h, _p(p, C.c_float)), "mlp_set")

    def get(self) -> np.ndarray:
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        self.ctx._ck(self.L.hsmc_gpu_mlp_get(self.h, _p(p, C.c_float)), "mlp_get")
        return p

    def forward(self, x) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.zeros(len(xx), dtype=np.float32)
        self.ctx._ck(self.L.hsmc_gpu_mlp_forward(self.h, _p(xx, C.c_float), len(xx),
                                                 _p(y, C.c_float)), "mlp_forward")
        return y

    def gradient(self, x, t) -> tuple:
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        tt = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = C.c_float(0)
        self.ctx._ck(self.L.hsmc_gpu_mlp_gradient(self.h, _p(xx, C.c_float), _p(tt, C.c_float),
                                                  len(xx), C.byref(loss), _p(g, C.c_float)),
                     "mlp_gradient")
        return loss.value, g

    def train(self, voxel: float, seed: int, sample: int, n_points: int, epochs: int = 10,
              batch: int = 4096, lr: float = 1e-3, beta: float = 20.0, r: int = 0) -> float:
        loss = C.c_float(0)
        self.ctx._ck(self.L.hsmc_gpu_mlp_train(self.h, r, float(voxel), int(seed), int(sample),
                                               int(n_points), int(epochs), int(batch),
                                               float(lr), float(beta), C.byref(loss)), "mlp_train")
        return loss.value

    def voxel_scores(self, voxel: float, r: int = 0) -> np.ndarray:
        nv = C.c_int64(0)
        self.ctx._ck(self.L.hsmc_gpu_voxel_scores(self.h, r, float(voxel), None, 0, C.byref(nv)),
                     "voxel_scores")
        out = np.zeros(nv.value, dtype=np.float32)
        self.ctx._ck(self.L.hsmc_gpu_voxel_scores(self.h, r, float(voxel), _p(out, C.c_float),
                                                  nv.value, C.byref(nv)), "voxel_scores")
        return out

    def forget(self) -> None:
        """Drop the cached proposal, so the next widom call scores again."""
        self.ctx._ck(self.L.hsmc_gpu_mlp_forget(self.h), "mlp_forget")

    def proposal_scores(self, voxel: float, r: int = 0) -> tuple:
        """The scores the learned proposal uses (-inf where no point of the
        voxel can be insertable) and the number of voxels scored."""
        nv, nf = C.c_int64(0), C.c_int64(0)
        self.ctx._ck(self.L.hsmc_gpu_proposal_scores(self.h, r, float(voxel), None, 0, C.byref(nv),
                                                     None), "proposal_scores")
        out = np.zeros(nv.value, dtype=np.float32)
        self.ctx._ck(self.L.hsmc_gpu_proposal_scores(self.h, r, float(voxel), _p(out, C.c_float),
                                                     nv.value, C.byref(nv), C.byref(nf)),
                     "proposal_scores")
        return out, nf.value

    def widom(self, voxel: float, mix: float, seed: int, sample: int, n: int, r: int = 0) -> dict:
        est = C.c_double(0)
        done = C.c_int64(0)
        hits = C.c_int64(0)
        self.ctx._ck(self.L.hsmc_gpu_widom_learned(self.h, r, float(voxel), float(mix), int(seed),
                                                   int(sample), int(n), C.byref(est),
                                                   C.byref(done), C.byref(hits)), "widom_learned")
        return dict(estimate=est.value, n=done.value, hits=hits.value)


def occupant_picks(seed: int, cell: int, sweep: int, n: int, count: int,
                   device: int = 0) -> np.ndarray:
    """The sweep kernel's occupant pick for attempts 0..count-1 of one cell
    among n (any n in [1, 2^32 - 1]), evaluated on the device."""
    out = np.zeros(count, dtype=np.uint32)
    st = load().hsmc_gpu_occupant_picks(device, int(seed), int(cell), int(sweep), int(n),
                                        int(count), _p(out, C.c_uint32))
    if st != HSMC_OK:
        raise HsmcError(st, "occupant_picks")
    return out


def point_descriptors(ctx: "Context", xyz, r: int = 0) -> np.ndarray:
    q = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(q), MLP_IN), dtype=np.float32)
    ctx._ck(ctx.L.hsmc_gpu_point_descriptors(ctx.h, r, _p(q, C.c_double), len(q),
                                             _p(out, C.c_float)), "point_descriptors")
    return out


def lattice(type_: int, n, rho: float):
    """FCC (type 2) or SC (type 1) starting lattice and box, as sim_box_init
    and part_init build them (sim_info.c:32-71, :125-156).  n is the number
    of unit cells per edge, or a tuple (cells_x, cells_y, cells_z)."""
    import math
    nx, ny, nz = (n, n, n) if isinstance(n, int) else tuple(int(v) for v in n)
    pc = 1 if type_ == 1 else 4
    a = math.pow(pc / rho, 1.0 / 3.0)
    pts = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if type_ == 1:
                    pts.append((i * a, j * a, k * a))
                else:
                    pts.append((i * a, j * a, k * a))
                    pts.append(((i + 0.5) * a, (j + 0.5) * a, k * a))
                    pts.append(((i + 0.5) * a, j * a, (k + 0.5) * a))
                    pts.append((i * a, (j + 0.5) * a, (k + 0.5) * a))
    conf = np.zeros((len(pts), 4))
    conf[:, 0] = np.arange(len(pts))
    conf[:, 1:] = np.array(pts)
    return conf, (nx * a, ny * a, nz * a)
