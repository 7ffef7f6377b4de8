"""ctypes view of the C host layer (libhsmc.so, include/hsmc.h): the global
input structure and the parser, for tests and tools that drive the layer
from Python.  The simulation itself runs through the ``hsmc`` executable or
the C entry points."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libhsmc.so")
CLI = os.path.join(LIB_DIR, "hsmc")


class Input(C.Structure):
    """struct input (read_input.h:4-69, plus the two GPU keywords)."""
    _fields_ = [
        ("rho", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("type", C.c_int), ("neigh_dr", C.c_double), ("neigh_max_part", C.c_int),
        ("dr_max", C.c_double), ("sweep_eq", C.c_int), ("sweep_stat", C.c_int),
        ("output_int", C.c_int), ("press", C.c_double), ("dv_max", C.c_double),
        ("opt_flag", C.c_int), ("opt_samples", C.c_int), ("opt_sweeps", C.c_int),
        ("opt_part_target", C.c_double), ("opt_vol_target", C.c_double),
        ("seed", C.c_ulong), ("cavity_pcav", C.c_double), ("cavity_maxdr", C.c_double),
        ("cavity_mindr", C.c_double), ("cavity_out_dr", C.c_double),
        ("cavity_sample_int", C.c_int), ("cluster_flag", C.c_int),
        ("cluster_moves_sweep", C.c_int), ("cluster_init_step", C.c_int),
        ("restart_read", C.c_int), ("restart_name", C.c_char * 100),
        ("restart_write", C.c_int), ("config_write", C.c_int), ("config_samples", C.c_int),
        ("pressv_dr", C.c_double), ("presst_xi_max", C.c_double), ("presst_dxi", C.c_double),
        ("pressv_sample_int", C.c_int), ("presst_sample_int", C.c_int),
        ("ql_order", C.c_int), ("ql_sample_int", C.c_int), ("ql_rmax", C.c_double),
        ("mu_insertions", C.c_int), ("mu_sample_int", C.c_int),
        ("rdf_sample_int", C.c_int), ("rdf_samples", C.c_int),
        ("rdf_dr", C.c_double), ("rdf_rmax", C.c_double),
        ("gpu_serial", C.c_int), ("gpu_device", C.c_int),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise ImportError(f"{HOST_LIB} is missing: run __graft_entry__.build()")
        C.CDLL(os.path.join(LIB_DIR, "libhsmc_gpu.so"), mode=C.RTLD_GLOBAL)
        _lib = C.CDLL(HOST_LIB)
        _lib.read_input_file.argtypes = [C.c_char_p]
        _lib.read_input_file.restype = None
    return _lib


def g_in() -> Input:
    return Input.in_dll(load(), "G_IN")


def read_input_file(path: str) -> Input:
    load().read_input_file(path.encode())
    return g_in()
