"""Post-processing of the output files: compressibility factor from the
contact value of g(r), thermodynamic-route pressure, Widom chemical
potential, NpT density (average, histogram, autocorrelation), averaged g(r)
with its blocking analysis and finite-size correction, the direct
correlation and bridge functions, and the blocking / jackknife error
estimates.  These restate the reference's analysis scripts (fedluc/HSMC
python/hsmc_pressure.py, hsmc_chem_pot.py, hsmc_stat.py, hsmc_density.py,
hsmc_rdf.py) in numpy; the reference's scripts are not used.
"""
from __future__ import annotations

import gzip
import math
from dataclasses import dataclass

import numpy as np


# ----------------------------------------------------------------- statistics

def blocking_std(data) -> np.ndarray:
    """Flyvbjerg-Petersen blocking (hsmc_stat.py:6-50): the series is cut to
    the largest power of two, then repeatedly averaged in pairs; level k gives
    sigma_k = std(ddof=0)/sqrt(n_k - 1) and its uncertainty
    sigma_k/sqrt(2(n_k - 1)).  Returns an array (levels, 2); levels =
    log2(n) - 1."""
    x = np.asarray(data, dtype=np.float64)
    n = len(x)
    if n < 4:
        raise ValueError("blocking needs at least 4 samples")
    p = n.bit_length() - 1
    x = x[: 2 ** p]
    out = np.zeros((p - 1, 2))
    for k in range(p - 1):
        m = len(x)
        s = np.std(x) / math.sqrt(m - 1)
        out[k, 0] = s
        out[k, 1] = s / math.sqrt(2 * (m - 1))
        x = 0.5 * (x[0::2] + x[1::2])
    return out


def blocking_error(data) -> float:
    """Plateau of the blocking curve: the largest sigma_k whose level still
    has at least 32 blocks (a conservative read of the plateau)."""
    sig = blocking_std(data)
    n = 2 ** (len(np.asarray(data)).bit_length() - 1)
    best = sig[0, 0]
    for k in range(len(sig)):
        if n // (2 ** k) >= 32:
            best = max(best, sig[k, 0])
    return float(best)


def jackknife(data) -> np.ndarray:
    """Leave-one-out mean and standard deviation (hsmc_stat.py:114-130)."""
    x = np.asarray(data, dtype=np.float64)
    n = len(x)
    xi = (x.sum() - x) / (n - 1)
    return np.array([xi.mean(), np.std(xi) * math.sqrt(n - 1)])


def linear_fit(x, y) -> tuple:
    """Least-squares line y = b0 + b1 x (the model of hsmc_pressure.py:214-228;
    the reference fits it with scipy curve_fit, whose optimum is this)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xm, ym = x.mean(), y.mean()
    b1 = ((x - xm) * (y - ym)).sum() / ((x - xm) ** 2).sum()
    return ym - b1 * xm, b1


# ------------------------------------------------------------------- readers

@dataclass
class Histogram:
    x: np.ndarray        # bin centres (or compressions)
    y: np.ndarray        # normalised histogram (or flags)
    vol: float
    n_part: int
    sweep: int = -1


def read_histogram_file(path: str) -> list:
    """press_virial.dat / press_thermo.dat: repeated samples of a 7-line
    header (count, volume, N on line 4) then count lines of two numbers
    (compute_press.c:277-359)."""
    with open(path) as f:
        lines = f.readlines()
    out, i = [], 0
    while i + 7 <= len(lines):
        nb, vol, n = lines[i + 3].split()
        nb = int(nb)
        rows = np.array([[float(v) for v in lines[i + 7 + j].split()] for j in range(nb)])
        out.append(Histogram(rows[:, 0], rows[:, 1], float(vol), int(n)))
        i += 7 + nb
    return out


def read_rdf_file(path: str) -> list:
    """rdf_%06d.dat.gz: header with (sweep, bins, volume, N) on line 4, then
    bins lines of (r, g) (compute_rdf.c:184-194)."""
    with gzip.open(path, "rt") as f:
        lines = f.readlines()
    out, i = [], 0
    while i + 7 <= len(lines):
        sw, nb, vol, n = lines[i + 3].split()
        nb = int(nb)
        rows = np.array([[float(v) for v in lines[i + 7 + j].split()] for j in range(nb)])
        out.append(Histogram(rows[:, 0], rows[:, 1], float(vol), int(n), int(sw)))
        i += 7 + nb
    return out


def read_columns(path: str) -> np.ndarray:
    """chem_pot.dat / density.dat: '#' comment lines, then numeric rows."""
    return np.loadtxt(path, comments="#", ndmin=2)


def read_config_file(path: str) -> list:
    """config_%06d.dat.gz snapshots (io_config.c:165-178): list of
    (sweep, box (3,), conf (N, 4))."""
    with gzip.open(path, "rt") as f:
        lines = f.readlines()
    out, i = [], 0
    while i < len(lines):
        sweep = int(lines[i + 1])
        n = int(lines[i + 3])
        box = tuple(float(lines[i + 5 + k]) for k in range(3))
        conf = np.array([[float(v) for v in lines[i + 9 + j].split()] for j in range(n)])
        out.append((sweep, box, conf))
        i += 9 + n
    return out


# --------------------------------------------------------------- estimators

def contact_value(h: Histogram) -> float:
    """g(sigma+) = b0 + b1 from a line through the contact bins
    (hsmc_pressure.py:25)."""
    b0, b1 = linear_fit(h.x, h.y)
    return b0 + b1


def compressibility_virial(samples: list) -> np.ndarray:
    """Z = 1 + (2 pi / 3) rho g(sigma+) per sample (hsmc_pressure.py:28-30)."""
    return np.array([1.0 + (2.0 * math.pi / 3.0) * (s.n_part / s.vol) * contact_value(s)
                     for s in samples])


def pressure_thermo(samples: list, densities=None) -> dict:
    """Thermodynamic route (hsmc_pressure.py:46-103): per block, the average
    no-overlap fraction f(xi) at each compression; -ln f fitted by a line
    whose slope over N is the excess pressure; beta P = rho (1 + p_ex)."""
    xi = samples[0].x
    f = np.mean([s.y for s in samples], axis=0)
    keep = f > 0
    b0, b1 = linear_fit(xi[keep], -np.log(f[keep]))
    n = samples[0].n_part
    p_ex = b1 / n
    rho = np.mean(densities) if densities is not None else n / samples[0].vol
    return {"p_ex": p_ex, "betaP": rho * (1.0 + p_ex), "Z": 1.0 + p_ex, "rho": rho}


def widom_mu(rows: np.ndarray, insertions: int) -> dict:
    """Excess chemical potential from chem_pot.dat rows (mu_s, hit fraction):
    the reference's average of mu_s over samples with hits
    (hsmc_chem_pot.py:14-30, :105-108), and the unbiased -ln(mean fraction)."""
    mu_s, frac = rows[:, 0], rows[:, 1]
    with_hits = frac > 0
    ref_style = float(np.mean(mu_s[with_hits])) if with_hits.any() else float("nan")
    mean_frac = float(np.mean(frac))
    unbiased = -math.log(mean_frac) if mean_frac > 0 else float("inf")
    return {"mu_reference_estimator": ref_style, "mu_ex": unbiased, "p_insert": mean_frac,
            "samples": len(rows), "insertions": insertions * len(rows)}


def average_rdf(samples: list) -> tuple:
    """Mean g(r) over samples with its standard error (hsmc_rdf.py:26-70)."""
    g = np.array([s.y for s in samples])
    return samples[0].x, g.mean(axis=0), g.std(axis=0, ddof=1) / math.sqrt(len(g))


def rdf_target_bins(r, target_r) -> list:
    """Index of the first bin whose r is >= each target distance, the rule of
    hsmc_rdf.py:100-111; a target beyond the last bin is an error."""
    r = np.asarray(r, dtype=np.float64)
    out = []
    for t in target_r:
        hit = np.nonzero(r >= t)[0]
        if len(hit) == 0:
            raise ValueError(f"target distance {t:.8f} lies beyond the last g(r) bin")
        out.append(int(hit[0]))
    return out


def rdf_blocking(samples: list, target_r) -> list:
    """Blocking analysis of g at chosen distances over all samples
    (hsmc_rdf.py:75-155): each target snaps to the first bin at or beyond
    it, and the sample series of that bin goes through blocking_std.
    Returns [(r_bin, blocking_std array (levels, 2)), ...]."""
    idx = rdf_target_bins(samples[0].x, target_r)
    g = np.array([s.y for s in samples])
    return [(float(samples[0].x[i]), blocking_std(g[:, i])) for i in idx]


def write_rdf_average(samples: list, out_dir: str = ".") -> str:
    """The averaged g(r) file of hsmc_rdf.py:59-70: two columns (r, mean g)
    at %.16e, named rdf_average_<samples>config.dat.  Returns the path."""
    import os
    r, g, _ = average_rdf(samples)
    path = os.path.join(out_dir, "rdf_average_%dconfig.dat" % len(samples))
    np.savetxt(path, np.column_stack([r, g]), fmt="%.16e")
    return path


def write_rdf_blocking(samples: list, target_r, source: str = ".",
                       path: str = "rdf_blocking_analysis.dat") -> str:
    """The blocking report of hsmc_rdf.py:123-155: a header (source, total
    samples, the snapped target distances at %.8f), then per target the
    blocking levels as 'level sigma uncertainty'.  Returns the path."""
    res = rdf_blocking(samples, target_r)
    with open(path, "w") as f:
        f.write("Blocking analysis of the radial distribution function data in " + source + "\n")
        f.write("Total number of samples: %d\n" % len(samples))
        f.write("Target positions:\n")
        for r, _ in res:
            f.write("%.8f\n" % r)
        for r, sig in res:
            f.write("###############################################\n")
            f.write("Blocking analysis for target distance %.8f\n" % r)
            f.write("Transformation, variance, variance uncertainty\n")
            for k in range(len(sig)):
                f.write("%d %.8f %.8f\n" % (k, sig[k, 0], sig[k, 1]))
    return path


def finite_size_correction(g, rho: float, n_part: int) -> np.ndarray:
    """g(r) of an N-particle canonical run scaled by 1 + rho / (N beta dP/drho),
    with beta dP/drho from the Pade equation of state below
    (hsmc_rdf.py:318-326, which writes the factor as 1 + 1/(N mu_T) with
    mu_T = (beta dP/drho)/rho).  The Lebowitz-Percus tail g_N -> 1 - S(0)/N,
    S(0) = 1/(beta dP/drho), would give 1 + S(0)/N, smaller by the factor
    rho; the reference's factor is kept."""
    iic = inverse_compressibility_ew(math.pi * rho / 6.0)
    return np.asarray(g, dtype=np.float64) * (1.0 + rho / (n_part * iic))


def _radial_grid(r, g):
    """Midpoint grid (k + 1/2) dr from 0 to the last g(r) bin, with h = g - 1
    on the bins and h = -1 inside the core below the first bin
    (hsmc_rdf.py:224-236)."""
    r = np.asarray(r, dtype=np.float64)
    dr = r[1] - r[0]
    n0 = int(round(r[0] / dr - 0.5))
    if n0 < 0 or abs((n0 + 0.5) * dr - r[0]) > 1e-6 * dr:
        raise ValueError("g(r) bins must be centred on the grid (k + 1/2) dr")
    rr = (np.arange(n0 + len(r)) + 0.5) * dr
    h = np.full(len(rr), -1.0)
    h[n0:] = np.asarray(g, dtype=np.float64) - 1.0
    return rr, h, dr


def direct_correlation(r, g, rho: float, n_part: int) -> dict:
    """Direct correlation function c(r) = h(r) - gamma(r) from the
    Ornstein-Zernike relation (hsmc_rdf.py:209-265).  g is first scaled by
    finite_size_correction.  The radial Fourier transform
    f(q) = (4 pi / q) int r f(r) sin(q r) dr is a midpoint sum, a type-II
    sine transform on the grid r_k = (k + 1/2) dr, q_j = (j + 1) pi/(n dr);
    gamma(q) = rho h(q)^2 / (1 + rho h(q)) goes back through the type-III
    transform, gamma(r) = (dq / (2 pi^2 r)) sum q gamma(q) sin(q r).
    Returns r, c, gamma, h, the corrected g and S(q) = 1 + rho h(q) on q."""
    from scipy.fft import dst
    gc = finite_size_correction(g, rho, n_part)
    rr, h, dr = _radial_grid(r, gc)
    n = len(rr)
    dq = math.pi / (n * dr)
    q = (np.arange(n) + 1.0) * dq
    hq = 4.0 * math.pi * dr / q * dst(rr * h, type=2) / 2.0
    gq = rho * hq * hq / (1.0 + rho * hq)
    gam = dq / (2.0 * math.pi ** 2 * rr) * dst(q * gq, type=3) / 2.0
    return {"r": rr, "c": h - gam, "gamma": gam, "h": h, "g": gc,
            "q": q, "S": 1.0 + rho * hq}


def bridge_function(r, g, rho: float, n_part: int, ln_y=None) -> dict:
    """Bridge function B(r) = ln y(r) - gamma(r) (hsmc_rdf.py:268-316), with
    ln y = ln g outside the core and NaN inside unless the cavity function
    ln y is given on the same grid."""
    d = direct_correlation(r, g, rho, n_part)
    rr = d["r"]
    if ln_y is None:
        ly = np.full(len(rr), np.nan)
        n0 = len(rr) - len(d["g"])
        with np.errstate(divide="ignore"):
            ly[n0:] = np.log(d["g"])
    else:
        ly = np.asarray(ln_y, dtype=np.float64)
        if ly.shape != rr.shape:
            raise ValueError("ln y must be given on the grid (k + 1/2) dr from 0")
    return {"r": rr, "B": ly - d["gamma"], "ln_y": ly, "c": d["c"], "gamma": d["gamma"]}


# ------------------------------------------------------------ density series

def density_histogram(series, bins: int = 100) -> tuple:
    """Histogram of the NpT density samples (hsmc_density.py:42-58)."""
    return np.histogram(np.asarray(series, dtype=np.float64).ravel(), bins=bins)


def acf(data, norm: bool = True) -> np.ndarray:
    """Autocorrelation by direct summation (hsmc_stat.py:69-88):
    c_k = sum_{j < n-k} (x_j - m)(x_{j+k} - m) / n, optionally over c_0."""
    x = np.asarray(data, dtype=np.float64)
    n = len(x)
    d = x - x.mean()
    c = np.array([np.dot(d[: n - k], d[k:]) / n for k in range(n)])
    return c / c[0] if norm else c


def acf_fft(data, norm: bool = True) -> np.ndarray:
    """The same autocorrelation through an FFT zero-padded to twice the next
    power of two, so the circular product has no wrap-around
    (hsmc_stat.py:91-103)."""
    x = np.asarray(data, dtype=np.float64)
    n = len(x)
    m = 1 << max(0, (n - 1).bit_length())
    f = np.fft.fft(x - x.mean(), n=2 * m)
    c = np.fft.ifft(f * np.conjugate(f))[:n].real / n
    return c / c[0] if norm else c


def density_acf(series, samples_block: int = -1, mode: str = "fft") -> np.ndarray:
    """Autocorrelation of the density series averaged over blocks of
    samples_block samples (the trailing remainder dropped; -1 or more than
    the series: one block), as hsmc_density.py:73-100 and :127-137."""
    x = np.asarray(series, dtype=np.float64).ravel()
    if samples_block < 0 or samples_block > len(x):
        samples_block = len(x)
    blocks = x[: len(x) - len(x) % samples_block].reshape(-1, samples_block)
    fn = {"fft": acf_fft, "dsum": acf}.get(mode)
    if fn is None:
        raise ValueError(f"unknown autocorrelation mode {mode!r}")
    return np.mean([fn(b) for b in blocks], axis=0)


# --------------------------------------------------------- closed-form anchors

# The Pade form of the hard-sphere fluid equation of state that the reference
# attributes to Erpenbeck and Wood, Z - 1 = eta P(eta) / Q(eta), with its
# coefficients (hsmc_rdf.py:333-339).  Its Taylor series reproduces the virial
# coefficients 4, 10 and 18.36; at eta = 0.25, 0.40 and 0.45 it lies 0.3%, 1.2%
# and 1.8% above the Monte Carlo pressures of De Miguel and Jackson.
_EW_P = (4.0, 0.890851, 0.8924486, 0.3430298)
_EW_Q = (1.0, -2.277287, 1.3262418)


def compressibility_ew(eta: float) -> float:
    """Z = beta P / rho of that Pade form."""
    p = _EW_P[0] + eta * (_EW_P[1] + eta * (_EW_P[2] + eta * _EW_P[3]))
    q = _EW_Q[0] + eta * (_EW_Q[1] + eta * _EW_Q[2])
    return 1.0 + eta * p / q


def inverse_compressibility_ew(eta: float) -> float:
    """beta dP/drho = d(rho Z)/drho = Z + eta dZ/deta for that Pade form
    (hsmc_rdf.py:329-352)."""
    p = _EW_P[0] + eta * (_EW_P[1] + eta * (_EW_P[2] + eta * _EW_P[3]))
    dp = _EW_P[1] + eta * (2.0 * _EW_P[2] + eta * 3.0 * _EW_P[3])
    q = _EW_Q[0] + eta * (_EW_Q[1] + eta * _EW_Q[2])
    dq = _EW_Q[1] + 2.0 * eta * _EW_Q[2]
    dz = p / q + eta * (dp * q - p * dq) / (q * q)   # d(eta P/Q)/d eta
    return 1.0 + eta * p / q + eta * dz


def carnahan_starling(rho: float) -> dict:
    """Carnahan-Starling Z, contact value and excess chemical potential."""
    eta = math.pi * rho / 6.0
    z = (1 + eta + eta ** 2 - eta ** 3) / (1 - eta) ** 3
    g = (1 - eta / 2) / (1 - eta) ** 3
    mu = (8 * eta - 9 * eta ** 2 + 3 * eta ** 3) / (1 - eta) ** 3
    return {"eta": eta, "Z": z, "g_contact": g, "mu_ex": mu}
