"""Re-binning of a samples file into a distribution (the PlotHistogram program of RecurGreen).

Mirrors src/plot_histogram.f90: read `&settings bintype, nbins, xmin, xmax /` (:55-59), read one real
per line from the data file, skipping lines that do not parse (read_real_array,
src/base_utils.f90:663-697), check the arguments (:64-71), build the bin lattice (:74-85), count on
the GPU (rg_histogram: last edge inclusive, out-of-range samples missing, NaN dropped,
src/base_utils.f90:491-516), add eps/10, normalise, take bin centres (:95-97), and write
`replot_N.distrib.csv` beside the data file with the data file's leading '%'/'#' lines copied
(save_histo :104-138, copy_preamble src/base_utils.f90:856-871).

Usage: python -m recurgreen_amd.replot plothisto_settings.nml path/to/samples.csv [--device D]
"""
from __future__ import annotations

import argparse
import math
import os
import re
import sys

import numpy as np

from . import binding, distrib
from .binding import RecurGreenError

BINTYPES = ("linear", "chebyshev")


def parse_namelist(text: str, group: str) -> dict:
    """Values of one namelist group (&group ... /) as Python values: quoted strings, integers,
    reals (E or D exponents), logicals; '!' starts a comment outside quotes."""
    lines = []
    for line in text.splitlines():
        out, q = [], None
        for ch in line:
            if q:
                out.append(ch)
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
                out.append(ch)
            elif ch == "!":
                break
            else:
                out.append(ch)
        lines.append("".join(out))
    body = "\n".join(lines)
    m = re.search(r"&" + re.escape(group) + r"\b(.*?)(?:^\s*/|/\s*$)", body, re.S | re.M | re.I)
    if not m:
        raise RecurGreenError(binding.RG_EINVAL, f"Namelist group '&{group}' not found")
    vals = {}
    for key, raw in re.findall(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\s/]+)", m.group(1)):
        raw = raw.strip()
        if raw[0] in "'\"":
            v = raw[1:-1]
        elif raw.lower() in (".true.", "t", ".t."):
            v = True
        elif raw.lower() in (".false.", "f", ".f."):
            v = False
        elif re.fullmatch(r"[+-]?\d+", raw):
            v = int(raw)
        else:
            v = float(raw.replace("d", "e").replace("D", "e"))
        vals[key.lower()] = v
    return vals


def check_settings(bintype: str, nbins: int, xmin: float, xmax: float) -> None:
    """src/plot_histogram.f90:65-72 and :84-85."""
    if nbins < 1:
        raise RecurGreenError(binding.RG_EINVAL, f"Invalid number of bins (nbins={nbins} must be positive), aborting...")
    if xmax <= xmin or math.isnan(xmin) or math.isnan(xmax):
        raise RecurGreenError(binding.RG_EINVAL, "Invalid binning range, aborting...")
    if bintype not in BINTYPES:
        raise RecurGreenError(binding.RG_EINVAL, f"Unknown binning type '{bintype}', aborting...")


def read_preamble(path: str) -> list[str]:
    """Leading lines that start with '%' or '#', each cut to 300 characters and right-trimmed."""
    out = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")[:300]
            if line[:1] in ("%", "#"):
                out.append(line.rstrip())
            else:
                break
    return out


def write_replot(path: str, rho: np.ndarray, preamble: list[str]) -> None:
    """The data file's preamble verbatim, 'tm, rho', then g0.16 rows (src/plot_histogram.f90:129-132)."""
    with open(path, "w") as f:
        for line in preamble:
            f.write(line + "\n")
        f.write("tm, rho\n")
        for tm, r in rho:
            f.write(f"{distrib.format_g016(tm)}, {distrib.format_g016(r)}\n")


def unique_path(directory: str, base: str, ext: str) -> str:
    """unique_filename (src/base_utils.f90:825-840): base + smallest i >= 1 + ext not yet present."""
    i = 1
    while os.path.exists(os.path.join(directory, f"{base}{i}{ext}")):
        i += 1
    return os.path.join(directory, f"{base}{i}{ext}")


def histogram_plan(device: int = 0) -> binding.Plan:
    """A minimal plan (two modes, two slices) whose only use is the device histogram."""
    return binding.Plan(dict(boundtype="dirichlet", nxdiso=2, wol=1.25), batch_max=1, device=device)


def plot_histo(samples: np.ndarray, bintype: str, nbins: int, xmin: float, xmax: float, plan, device: int = 0):
    """(rho_data[nbins, 2], missing): plot_histo_from_file (src/plot_histogram.f90:52-98) on the
    device histogram.  Samples are copied to device memory once and counted there."""
    check_settings(bintype, nbins, xmin, xmax)
    edges = distrib.bins(bintype, nbins, xmin, xmax)
    counts, missing = device_counts(plan, samples, edges, device)
    return distrib.rho_from_counts(counts, edges), missing


def device_counts(plan, samples: np.ndarray, edges: np.ndarray, device: int = 0):
    """(counts[nbins], missing) from the device histogram kernel, samples staged in device memory."""
    import torch

    dev = torch.device("cuda", device)
    s = torch.as_tensor(np.ascontiguousarray(samples, dtype=np.float64)).to(dev)
    torch.cuda.synchronize(dev)
    counts, missing, _dropped = plan.histogram(s, edges, on_device=True)
    return counts, missing


def replot(config_path: str, data_path: str, plan=None, device: int = 0, log=print) -> str:
    """Write replot_N.distrib.csv beside `data_path`; returns its path."""
    for p in (config_path, data_path):
        if not os.path.exists(p):
            raise RecurGreenError(binding.RG_EINVAL, f"File not found: {p}")
    st = parse_namelist(open(config_path).read(), "settings")
    bintype, nbins = str(st.get("bintype", "")), int(st.get("nbins", 0))
    xmin, xmax = float(st.get("xmin", math.nan)), float(st.get("xmax", math.nan))
    samples = distrib.read_real_array(data_path)
    log(f"[INFO] Found {samples.size} samples in the data file...")
    check_settings(bintype, nbins, xmin, xmax)
    own = plan is None
    if own:
        plan = histogram_plan(device)
    try:
        rho, missing = plot_histo(samples, bintype, nbins, xmin, xmax, plan, device)
    finally:
        if own:
            plan.close()
    if missing:
        e = distrib.bins(bintype, nbins, xmin, xmax)
        log(f"[WARN] Some samples are out of the bounds [{e[0]:.6f}, {e[-1]:.6f}].")
        log(f"[WARN] There are {int(missing)} missing samples in the histogram ({100.0 * missing / samples.size:.2f}%), "
            "continuing anyway...")
    out = unique_path(os.path.dirname(os.path.abspath(data_path)), "replot_", ".distrib.csv")
    write_replot(out, rho, read_preamble(data_path))
    log(f"[INFO] Replot written in file: {out}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Re-bin a samples file into replot_N.distrib.csv (PlotHistogram).")
    ap.add_argument("settings")
    ap.add_argument("data")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    try:
        replot(a.settings, a.data, device=a.device)
    except RecurGreenError as e:
        print(f"[ERROR] {e}", file=sys.stderr)
        print("RecurGreen: Invalid arguments, aborting...", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
