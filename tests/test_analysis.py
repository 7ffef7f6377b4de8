"""Post-processing restatements against closed forms."""
import math

import numpy as np

from hsmc_amd import analysis as A


def test_linear_fit_is_least_squares():
    rng = np.random.default_rng(1)
    x = rng.random(20)
    y = 3.0 - 2.0 * x + 0.01 * rng.standard_normal(20)
    b0, b1 = A.linear_fit(x, y)
    p1, p0 = np.polyfit(x, y, 1)
    assert abs(b0 - p0) < 1e-12 and abs(b1 - p1) < 1e-12
    # exact line is recovered exactly
    b0, b1 = A.linear_fit([1.005, 1.015, 1.025], [5.0, 4.9, 4.8])
    assert abs(b1 + 10.0) < 1e-9 and abs((b0 + b1) - 5.05) < 1e-9


def test_jackknife_of_mean_closed_form():
    x = np.random.default_rng(2).standard_normal(101)
    m, s = A.jackknife(x)
    assert abs(m - x.mean()) < 1e-12
    assert abs(s - x.std(ddof=1) / math.sqrt(len(x))) < 1e-12


def test_blocking_levels():
    x = np.random.default_rng(3).standard_normal(1 << 12)
    b = A.blocking_std(x)
    assert b.shape == (11, 2)
    assert abs(b[0, 0] - x.std() / math.sqrt(len(x) - 1)) < 1e-15
    y = 0.5 * (x[0::2] + x[1::2])
    assert abs(b[1, 0] - y.std() / math.sqrt(len(y) - 1)) < 1e-15
    # uncorrelated data: the plateau is the naive standard error
    assert abs(A.blocking_error(x) / (x.std() / math.sqrt(len(x))) - 1) < 0.35
    # a correlated series (block-constant) has a larger blocked error
    z = np.repeat(np.random.default_rng(4).standard_normal(1 << 8), 16)
    assert A.blocking_error(z) > 2.5 * z.std() / math.sqrt(len(z))


def test_carnahan_starling_anchors():
    cs = A.carnahan_starling(0.85)
    assert abs(cs["Z"] - 9.099) < 2e-3
    assert abs(cs["g_contact"] - 4.549) < 2e-3
    assert abs(cs["mu_ex"] - 11.950) < 2e-3
    cs = A.carnahan_starling(0.5)
    assert abs(cs["Z"] - 3.262) < 2e-3 and abs(cs["mu_ex"] - 3.807) < 2e-3
    # Z = 1 + 4 eta g(sigma+) (virial theorem for hard spheres)
    eta = cs["eta"]
    assert abs(cs["Z"] - (1 + 4 * eta * cs["g_contact"])) < 1e-12


def test_histogram_file_round_trip(tmp_path):
    p = tmp_path / "press_virial.dat"
    with open(p, "w") as f:
        for s in range(3):
            f.write("######################################\n# Bins, volume, number of particles\n")
            f.write("######################################\n")
            f.write("%d %.8e %d\n" % (5, 1000.0, 500))
            f.write("###############################\n# rr, rdf\n###############################\n")
            for i in range(5):
                f.write("%.8e %.8e\n" % (1.005 + 0.01 * i, 2.2 - 0.1 * i + s))
    h = A.read_histogram_file(str(p))
    assert len(h) == 3 and h[0].n_part == 500 and h[0].vol == 1000.0
    # y = 2.2 + s - 10 (x - 1.005): the line's value at r = 1 (b0 + b1) is 2.25 + s
    for s_, hs in enumerate(h):
        assert abs(A.contact_value(hs) - (2.25 + s_)) < 1e-6
    z = A.compressibility_virial(h)
    assert abs(z[0] - (1 + 2 * math.pi / 3 * 0.5 * A.contact_value(h[0]))) < 1e-12


def test_widom_estimators():
    rows = np.array([[-math.log(0.5), 0.5], [0.0, 0.0], [-math.log(0.25), 0.25]])
    w = A.widom_mu(rows, 100)
    assert abs(w["mu_reference_estimator"] - 0.5 * (math.log(2) + math.log(4))) < 1e-12
    assert abs(w["mu_ex"] + math.log(0.25)) < 1e-12


def test_read_columns_single_and_multi_column(tmp_path):
    """density.dat holds one column, chem_pot.dat two; both read as (rows, cols)."""
    (tmp_path / "d.dat").write_text("####\n# Density\n####\n7.5e-01\n7.6e-01\n7.4e-01\n")
    d = A.read_columns(str(tmp_path / "d.dat"))
    assert d.shape == (3, 1) and d[1, 0] == 0.76
    (tmp_path / "m.dat").write_text("# mu, frac\n1.0 0.5\n")
    m = A.read_columns(str(tmp_path / "m.dat"))
    assert m.shape == (1, 2) and m[0, 1] == 0.5
