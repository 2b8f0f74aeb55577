"""Histogram chain of the tdistrib task and its on-disk formats.

Restates src/main.f90:223-251 (prepare_distrib), src/base_utils.f90:468-570 (check_histogram,
histogram, normalize_histogram, moving_average), src/plot_histogram.f90:77-98 (bin lattices)
and src/base_utils.f90:663-697 (read_real_array).  The counting itself runs on the GPU
(rg_histogram); normalisation is done here in FP64 with the reference's order of operations.
"""
from __future__ import annotations

import math
import re

import numpy as np

EPS_FLOOR = np.finfo(np.float64).eps / 10.0  # src/main.f90:243


class HistogramError(ValueError):
    pass


def bins(bintype: str, nbins: int, xmin: float = 0.0, xmax: float = 1.0) -> np.ndarray:
    """Bin edges.  'chebyshev': xmin + (xmax-xmin)*(1-cos(pi*(i-1)/nbins))/2 (src/main.f90:233,
    src/plot_histogram.f90:83); 'linear': xmin + (xmax-xmin)*(i-1)/nbins (:81)."""
    if nbins < 1:
        raise HistogramError(f"Invalid number of bins (nbins={nbins} must be positive), aborting...")
    if not (xmax > xmin):
        raise HistogramError("Invalid binning range, aborting...")
    e = np.empty(nbins + 1)
    for i in range(nbins + 1):
        if bintype == "linear":
            e[i] = xmin + (xmax - xmin) * i / nbins
        elif bintype == "chebyshev":
            e[i] = xmin + (xmax - xmin) * (1.0 - math.cos(math.pi * i / nbins)) / 2
        else:
            raise HistogramError(f"Unknown binning type '{bintype}', aborting...")
    return e


def check_histogram(edges: np.ndarray, nbins: int) -> None:
    """src/base_utils.f90:468-482."""
    if len(edges) - 1 != nbins:
        raise HistogramError("Invalid histogram length, aborting...")
    if np.any(edges[1:] < edges[:-1]):
        raise HistogramError("Unsorted array of bins, aborting...")


def normalize_histogram(edges: np.ndarray, histo: np.ndarray) -> np.ndarray:
    """histo_i / (sum(histo) * (b_{i+1} - b_i)), src/base_utils.f90:524-537 (sequential sum)."""
    check_histogram(edges, len(histo))
    npoints = 0.0
    for h in histo:
        npoints += float(h)
    return np.array([histo[i] / (npoints * (edges[i + 1] - edges[i])) for i in range(len(histo))])


def moving_average(signal: np.ndarray, n: int) -> np.ndarray:
    """src/base_utils.f90:543-570."""
    if n <= 1:
        raise HistogramError(f"Invalid moving average parameter {n}, expected >1, aborting...")
    out = np.empty(len(signal) - n + 1)
    for i in range(len(out)):
        s = 0.0
        for j in range(n):
            s += float(signal[i + j])
        out[i] = s / n
    return out


def rho_from_counts(counts: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """(tm, rho) rows from integer counts: +eps/10 floor, normalisation, bin centres
    (src/main.f90:243-246)."""
    h = np.asarray(counts, dtype=np.float64) + EPS_FLOOR
    rho = normalize_histogram(edges, h)
    cent = moving_average(edges, 2)
    return np.stack([cent, rho], axis=1)


def prepare_distrib(plan, tallvals, nbins: int, bintype: str = "chebyshev", xmin: float = 0.0,
                    xmax: float = 1.0):
    """prepare_distrib (src/main.f90:223-251) with the counting on the GPU.

    Returns (rho_data[nbins, 2], missing, dropped_nan)."""
    e = bins(bintype, nbins, xmin, xmax)
    counts, missing, dropped = plan.histogram(tallvals, e)
    return rho_from_counts(counts, e), missing, dropped


# -- on-disk formats ---------------------------------------------------------------------------
_FORTRAN_REAL = re.compile(r"^[+-]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][+-]?\d+)?$")


def _parse_fortran_real(tok: str):
    t = tok.strip()
    low = t.lower()
    if low in ("nan", "+nan", "-nan"):
        return float("nan")
    if low in ("inf", "+inf", "infinity", "+infinity"):
        return float("inf")
    if low in ("-inf", "-infinity"):
        return float("-inf")
    if _FORTRAN_REAL.match(t):
        return float(t.replace("d", "e").replace("D", "e"))
    return None


def read_real_array(path: str) -> np.ndarray:
    """One real per record, list-directed like src/base_utils.f90:663-697: the first value of a
    record is read, records that do not start with a valid real are skipped, blank records are
    passed over."""
    vals = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            tok = re.split(r"[,\s/]+", s, maxsplit=1)[0]
            v = _parse_fortran_real(tok)
            if v is not None:
                vals.append(v)
    return np.array(vals, dtype=np.float64)


def format_g016(x: float) -> str:
    """Fortran g0.16 for the values the distrib files hold (src/base_utils.f90:639-657)."""
    if x != x:
        return "NaN"
    ax = abs(x)
    if ax != 0.0 and (ax < 0.1 or ax >= 1e16):
        m, ex = f"{x:.15E}".split("E")
        mant = float(m) / 10.0
        e = int(ex) + 1
        s = f"{mant:.16f}".replace("0.", ".", 1) if mant >= 0 else "-" + f"{-mant:.16f}".replace("0.", ".", 1)
        s = "0" + s if s.startswith(".") else s.replace("-.", "-0.")
        return f"{s}E{e:+03d}"
    digits_before = len(str(int(ax))) if ax >= 1 else 0
    return f"{x:.{16 - digits_before}f}"


def save_distrib(path: str, rho_data: np.ndarray, header_lines=()) -> None:
    """result_N.distrib.csv layout of src/main.f90:257-294: '%%' header lines, 'tm, rho', rows."""
    with open(path, "w") as f:
        for h in header_lines:
            f.write(f"%% {h}\n" if not h.startswith("%%") else h + "\n")
        f.write("tm, rho\n")
        for tm, rho in rho_data:
            f.write(f"{format_g016(tm)}, {format_g016(rho)}\n")


def load_distrib(path: str) -> np.ndarray:
    rows = []
    with open(path) as f:
        for line in f:
            if not line.strip() or line.startswith("%") or line.startswith("tm"):
                continue
            a, b = line.split(",")
            rows.append((float(a), float(b)))
    return np.array(rows)
