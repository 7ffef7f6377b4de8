"""C host layer without a device: the keyword parser against the
reference's own parser (golden dumps), and the command line's error paths
(read_input.c:477-499, exec.c:28-64)."""
import os
import subprocess

import numpy as np
import pytest

from hsmc_amd import host_layer as H
from tests.conftest import ROOT

FIELDS = ["rho", "nx", "ny", "nz", "type", "neigh_dr", "neigh_max_part", "dr_max", "sweep_eq",
          "sweep_stat", "output_int", "press", "dv_max", "opt_flag", "seed", "config_write",
          "config_samples", "pressv_dr", "pressv_sample_int", "presst_dxi", "presst_xi_max",
          "presst_sample_int", "mu_insertions", "mu_sample_int", "rdf_dr", "rdf_rmax",
          "rdf_sample_int", "rdf_samples"]


@pytest.mark.parametrize("name", ["c1", "npt", "defaults"])
def test_parser_matches_reference_parser(golden, name):
    gi = H.read_input_file(os.path.join(ROOT, "tests", "golden", f"parse_{name}.in"))
    got = np.array([float(getattr(gi, f)) for f in FIELDS])
    assert np.array_equal(got, golden[f"parse/{name}"])


def _cli(tmp_path, text, *args):
    p = tmp_path / "in.dat"
    p.write_text(text)
    return subprocess.run([H.CLI, "-i", str(p), *args], cwd=tmp_path, capture_output=True,
                          text=True, timeout=60)


def test_cli_example_parses(tmp_path):
    r = subprocess.run([H.CLI, "-e"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "rho 0.5" in r.stdout
    p = tmp_path / "ex.dat"
    p.write_text(r.stdout)
    gi = H.read_input_file(str(p))
    assert gi.rho == 0.5 and gi.nx == 5 and gi.type == 2 and gi.pressv_sample_int == 10
    assert gi.mu_insertions == 100 and gi.rdf_samples == 100 and gi.opt_flag == 0


def test_cli_unknown_key(tmp_path):
    r = _cli(tmp_path, "rho 0.5\nbogus 1\n")
    assert r.returncode != 0
    assert "Unknown key" in r.stdout and "bogus" in r.stdout


def test_cli_missing_value(tmp_path):
    r = _cli(tmp_path, "neigh_list 1.05\n")
    assert r.returncode != 0 and "Missing value to key" in r.stdout


def test_cli_restart_name_too_long(tmp_path):
    r = _cli(tmp_path, "restart_read 1 " + "x" * 120 + "\n")
    assert r.returncode != 0 and "too long" in r.stdout


def test_cli_missing_input_file(tmp_path):
    r = subprocess.run([H.CLI, "-i", str(tmp_path / "nope.dat")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "Could not open file" in r.stdout


def test_extension_keywords(tmp_path):
    p = tmp_path / "x.dat"
    p.write_text("gpu_sweep serial\ngpu_device 0\nseed 7\n")
    gi = H.read_input_file(str(p))
    assert gi.gpu_serial == 1 and gi.gpu_device == 0 and gi.seed == 7
    p.write_text("gpu_sweep checkerboard\n")
    assert H.read_input_file(str(p)).gpu_serial == 0


CALLER_DIR = os.path.join(ROOT, "tests", "c_caller")


def build_caller(out_dir):
    """Compile the reference-style caller against include/hsmc.h and link it
    with libhsmc.so (every symbol it names must resolve)."""
    subprocess.run(["make", "-s", "-C", CALLER_DIR, f"OUT={out_dir}"], check=True, timeout=120)
    return os.path.join(str(out_dir), "caller")


def test_reference_style_caller_links(tmp_path):
    exe = build_caller(tmp_path)
    assert os.access(exe, os.X_OK)
    dyn = subprocess.run(["ldd", exe], capture_output=True, text=True, timeout=60).stdout
    assert "libhsmc.so" in dyn and "libhsmc_gpu.so" in dyn
