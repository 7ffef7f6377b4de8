"""Post-processing of the output files: compressibility factor from the
contact value of g(r), thermodynamic-route pressure, Widom chemical
potential, NpT density, averaged g(r), and the blocking / jackknife error
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


# --------------------------------------------------------- closed-form anchors

def carnahan_starling(rho: float) -> dict:
    """Carnahan-Starling Z, contact value and excess chemical potential."""
    eta = math.pi * rho / 6.0
    z = (1 + eta + eta ** 2 - eta ** 3) / (1 - eta) ** 3
    g = (1 - eta / 2) / (1 - eta) ** 3
    mu = (8 * eta - 9 * eta ** 2 + 3 * eta ** 3) / (1 - eta) ** 3
    return {"eta": eta, "Z": z, "g_contact": g, "mu_ex": mu}
