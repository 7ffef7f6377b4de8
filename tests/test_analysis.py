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


def test_equation_of_state_virial_coefficients():
    """The Pade Z - 1 = eta P/Q expands to 4 eta + 10 eta^2 + 18.36 eta^3 + ...,
    the exact hard-sphere virial coefficients B2-B4 (a wrong P or Q
    coefficient moves one of them)."""
    for eta in (1e-4, 2e-4):
        b2 = (A.compressibility_ew(eta) - 1.0) / eta
        assert abs(b2 - 4.0 - 10.0 * eta) < 20.0 * eta * eta
    e1, e2 = 1e-3, 2e-3
    c1 = ((A.compressibility_ew(e1) - 1) / e1 - 4) / e1
    c2 = ((A.compressibility_ew(e2) - 1) / e2 - 4) / e2
    b4 = (c2 - c1) / (e2 - e1)   # slope of (Z - 1 - 4 eta)/eta^2 near 0
    assert abs(c1 - 10.0) < 0.03 and abs(b4 - 18.365) < 0.2


def test_equation_of_state_against_simulation_and_cs():
    """Above De Miguel and Jackson's Monte Carlo pressures (the table in the
    reference's README) by 0.3% at eta = 0.25 growing to 1.8% at 0.45, and
    within 1.5% of Carnahan-Starling."""
    for eta, bp, _ in [(0.25, 1.4688, 0), (0.30, 2.2784, 0), (0.35, 3.4841, 0),
                       (0.40, 5.288, 0), (0.45, 8.030, 0)]:
        rho = 6.0 * eta / math.pi
        z = A.compressibility_ew(eta)
        assert 0.0 < z * rho / bp - 1.0 < 0.02
        assert abs(z / A.carnahan_starling(rho)["Z"] - 1.0) < 0.015


def test_inverse_compressibility_is_the_derivative():
    """beta dP/drho against a central difference of eta Z(eta), and against
    the Carnahan-Starling closed form (1+4e+4e^2-4e^3+e^4)/(1-e)^4."""
    for eta in (0.05, 0.2, 0.35, 0.45):
        h = 1e-6
        fd = ((eta + h) * A.compressibility_ew(eta + h)
              - (eta - h) * A.compressibility_ew(eta - h)) / (2 * h)
        iic = A.inverse_compressibility_ew(eta)
        assert abs(iic / fd - 1.0) < 1e-8
        cs = (1 + 4 * eta + 4 * eta ** 2 - 4 * eta ** 3 + eta ** 4) / (1 - eta) ** 4
        assert abs(iic / cs - 1.0) < 0.03
    assert abs(A.inverse_compressibility_ew(0.0) - 1.0) < 1e-15


def test_finite_size_correction():
    g = np.linspace(0.9, 1.1, 7)
    rho, n = 0.8, 500
    out = A.finite_size_correction(g, rho, n)
    eta = math.pi * rho / 6
    cs = (1 + 4 * eta + 4 * eta ** 2 - 4 * eta ** 3 + eta ** 4) / (1 - eta) ** 4
    # the factor is 1 + rho/(N beta dP/drho): check its size with the CS value
    ratio = (out / g - 1.0) * n * cs / rho
    assert np.all(np.abs(ratio - 1.0) < 0.03)
    assert np.allclose(out / g, out[0] / g[0], rtol=0, atol=1e-15)   # one factor for all r
    assert np.allclose(A.finite_size_correction(g, rho, 10 ** 12), g, rtol=1e-12)
    assert np.all(g == np.linspace(0.9, 1.1, 7))   # input left unchanged


def test_autocorrelation_closed_form_and_fft():
    n = 64
    x = np.array([(-1.0) ** j for j in range(n)])
    c = A.acf(x)
    k = np.arange(n)
    assert np.allclose(c, (-1.0) ** k * (n - k) / n, atol=1e-14)
    assert np.allclose(A.acf(x, norm=False), (-1.0) ** k * (n - k) / n, atol=1e-14)
    for m in (1000, 1024, 37):
        y = np.cumsum(np.random.default_rng(m).standard_normal(m))
        assert np.allclose(A.acf_fft(y), A.acf(y), atol=1e-10)
        assert np.allclose(A.acf_fft(y, norm=False), A.acf(y, norm=False), atol=1e-8)


def test_density_acf_blocks():
    rng = np.random.default_rng(7)
    b = rng.standard_normal(50)
    x = np.concatenate([b, b, rng.standard_normal(7)])   # 7 left over with blocks of 50
    assert np.allclose(A.density_acf(x, 50), A.acf(b), atol=1e-12)
    assert np.allclose(A.density_acf(x, 50, mode="dsum"), A.acf(b), atol=1e-12)
    assert np.allclose(A.density_acf(x), A.acf(x), atol=1e-10)
    assert np.allclose(A.density_acf(x, 10 ** 6), A.acf(x), atol=1e-10)
    try:
        A.density_acf(x, 10, mode="nope")
        raise AssertionError("an unknown mode must raise")
    except ValueError:
        pass
    h, edges = A.density_histogram(x, bins=10)
    assert h.sum() == len(x) and len(edges) == 11


def _rdf_samples(ns=64, nb=400, seed=0):
    r = 1.005 + 0.01 * np.arange(nb)
    rng = np.random.default_rng(seed)
    return [A.Histogram(r, 1.0 + 0.1 * rng.standard_normal(nb), 1000.0, 500, s)
            for s in range(ns)]


def test_rdf_blocking_at_target_distances(tmp_path):
    s = _rdf_samples()
    # first bin at or beyond the target: 1.0 -> bin 0, 1.005 -> 0, 1.0051 -> 1, 2.0 -> 100
    assert A.rdf_target_bins(s[0].x, [1.0, 1.005, 1.0051, 2.0]) == [0, 0, 1, 100]
    try:
        A.rdf_target_bins(s[0].x, [5.0])
        raise AssertionError("a target beyond the last bin must raise")
    except ValueError:
        pass
    res = A.rdf_blocking(s, [1.5, 3.0])
    g = np.array([h.y for h in s])
    for (r, sig), i in zip(res, [50, 200]):
        assert r == s[0].x[i]
        assert np.array_equal(sig, A.blocking_std(g[:, i]))
    p = A.write_rdf_blocking(s, [1.5, 3.0], source="runs", path=str(tmp_path / "b.dat"))
    lines = open(p).read().splitlines()
    assert lines[1] == "Total number of samples: 64" and lines[3] == "1.50500000"
    assert lines[6] == "Blocking analysis for target distance 1.50500000"
    lv = [ln for ln in lines if ln.startswith("0 ")][0].split()
    assert abs(float(lv[1]) - res[0][1][0, 0]) < 1e-8
    a = np.loadtxt(A.write_rdf_average(s, str(tmp_path)))
    assert a.shape == (400, 2) and np.allclose(a[:, 1], g.mean(axis=0), atol=1e-15)
    assert (tmp_path / "rdf_average_64config.dat").exists()


def test_direct_correlation_low_density_limit():
    """With g = 1 outside the core at rho -> 0: h = c = -1 inside, 0 outside,
    and gamma = rho (f*f)(r) to first order, the overlap volume of two unit
    balls, (pi/12)(4 + r)(2 - r)^2 for r < 2 (a closed form independent of
    the transform)."""
    r = 1.0 + (np.arange(700) + 0.5) * 0.01
    rho = 1e-4
    d = A.direct_correlation(r, np.ones_like(r), rho, 10 ** 15)
    rr = d["r"]
    assert len(rr) == 800 and abs(rr[0] - 0.005) < 1e-12 and abs(rr[100] - r[0]) < 1e-12
    v = np.where(rr < 2.0, math.pi / 12.0 * (4.0 + rr) * (2.0 - rr) ** 2, 0.0)
    far = rr > 0.1
    assert np.max(np.abs(d["gamma"][far] / rho - v[far])) < 0.02   # V(0) = 4.19
    assert np.max(np.abs(d["c"][rr < 1] + 1.0)) < 1e-3
    assert np.max(np.abs(d["c"][rr > 1])) < 1e-3
    b = A.bridge_function(r, np.ones_like(r), rho, 10 ** 15)
    assert np.all(np.isnan(b["B"][rr < 1]))
    assert np.allclose(b["B"][rr > 1], -d["gamma"][rr > 1], atol=1e-15)
    ly = np.zeros_like(rr)
    b2 = A.bridge_function(r, np.ones_like(r), rho, 10 ** 15, ln_y=ly)
    assert np.allclose(b2["B"], -d["gamma"], atol=1e-15)


def test_direct_correlation_satisfies_ornstein_zernike():
    """At rho = 0.3 with a structured g (S(q) > 0 throughout), the returned c and h satisfy
    h(q) = c(q) + rho c(q) h(q) under a radial transform computed by a direct
    sine sum (not the library's sine transform), on the transform's q grid."""
    r = 1.0 + (np.arange(400) + 0.5) * 0.01
    g = 1.0 + 2.5 * np.exp(-(r - 1.0) / 0.5) * np.cos(6.0 * (r - 1.0)) / r
    rho, n = 0.3, 4000
    d = A.direct_correlation(r, g, rho, n)
    assert np.allclose(d["g"], A.finite_size_correction(g, rho, n), rtol=0, atol=0)
    assert np.all(d["S"] > 0.0)
    rr, dr = d["r"], d["r"][1] - d["r"][0]
    q = d["q"][:200]
    sin = np.sin(np.outer(q, rr))
    ft = lambda f: 4.0 * math.pi * dr / q * (sin @ (rr * f))
    hq, cq = ft(d["h"]), ft(d["c"])
    assert np.allclose(hq, 1.0 / rho * (d["S"][:200] - 1.0), rtol=0, atol=1e-9)
    assert np.max(np.abs(hq - cq - rho * cq * hq)) < 1e-9
    assert np.allclose(d["h"][100:], d["g"] - 1.0, atol=0) and np.all(d["h"][:100] == -1.0)
    try:
        A.direct_correlation(r + 0.003, g, rho, n)
        raise AssertionError("bins off the midpoint grid must raise")
    except ValueError:
        pass
