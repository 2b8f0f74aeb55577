"""Transmission-eigenstate profiles inside the medium (tasks `tprofile` and `tprofile_group`).

compute_tprofile (src/waveguide.f90:843-907) sums, over the transmission eigenvalues T_v in
[tmin, tmax], the intensity |G(x;1) s U(:,v)|^2 of the eigenstate at every slice x and multiplies
by 4/nt; compute_tprofile_group (:920-997) does the same for several windows |T - tm| <= dt of the
aperture-filtered matrix and also returns tavg = sum(T)/neigv.  The device pass rg_tprofile
returns, for a batch of realizations, the eigenvalues and every per-eigenvector intensity (one
transmission sweep that keeps the Green functions of all slices, a Jacobi decomposition of t^H t,
and the back substitution of the block-tridiagonal solve); the window sums below are the
reference's steps 7 on those arrays.  The task drivers average the profiles of the realizations
with nt > 0 (src/main.f90:466-483).
"""
from __future__ import annotations

import datetime
import os
import time

import numpy as np

from . import binding
from .binding import RecurGreenError
from .distrib import format_g016
from .replot import unique_path
from .tpattern import parameter_lines


def check_window(tmin: float, tmax: float) -> None:
    """src/waveguide.f90:855-868."""
    if tmin < 0.0 or tmin > 1.0:
        raise RecurGreenError(binding.RG_EINVAL, f"Invalid Tmin, received Tmin={tmin:.3g}, expected in [0, 1], aborting...")
    if tmax < 0.0 or tmax > 1.0:
        raise RecurGreenError(binding.RG_EINVAL, f"Invalid Tmax, received Tmax={tmax:.3g}, expected in [0, 1], aborting...")
    if tmin >= tmax:
        raise RecurGreenError(binding.RG_EINVAL, "Invalid range (Tmin >= Tmax), aborting...")


def check_group(tm, dt) -> None:
    """src/waveguide.f90:938-949."""
    tm, dt = np.asarray(tm, float), np.asarray(dt, float)
    if tm.shape != dt.shape:
        raise RecurGreenError(binding.RG_EINVAL, "tm and dt must have the same length")
    if np.any(tm < 0.0) or np.any(tm > 1.0):
        raise RecurGreenError(binding.RG_EINVAL, "Invalid transmission value, expected in [0, 1], aborting...")
    if np.any(dt <= 0.0) or np.any(dt > 0.5):
        raise RecurGreenError(binding.RG_EINVAL, "Invalid transmission half-width, expected 0 < dt <= 1/2, aborting...")


def window_profile(tvals: np.ndarray, intensity: np.ndarray, tmin: float, tmax: float):
    """(profile[nx], nt) of one realization (src/waveguide.f90:887-903): tvals (nv,), intensity (nx, nv)."""
    sel = (tmin <= tvals) & (tvals <= tmax)
    nt = int(sel.sum())
    prof = intensity[:, sel].sum(axis=1)
    if nt:
        prof = prof * (4.0 / nt)
    return prof, nt


def itx_window(tvals: np.ndarray, intensity: np.ndarray, tmin: float, tmax: float, wol: float):
    """itx[nx] of one realization (src/waveguide.f90:808-821): the same window sum as
    window_profile, normalized by 2/(pi wol |tmax - tmin|) instead of 4/nt (the reference's older
    convention, I_T(x) = sum_n |psi_n|^2 delta(T - T_n) with a rectangular delta)."""
    sel = (tmin <= tvals) & (tvals <= tmax)
    return intensity[:, sel].sum(axis=1) * (2.0 / (np.pi * wol * abs(tmax - tmin)))


def group_profiles(tvals: np.ndarray, intensity: np.ndarray, tm, dt):
    """(tprofile[nx, nprofile], nt[nprofile], tavg) of one realization (src/waveguide.f90:973-990)."""
    tm, dt = np.asarray(tm, float), np.asarray(dt, float)
    prof = np.zeros((intensity.shape[0], tm.size))
    nt = np.zeros(tm.size, dtype=int)
    for ip in range(tm.size):
        sel = np.abs(tvals - tm[ip]) <= dt[ip]
        nt[ip] = int(sel.sum())
        prof[:, ip] = intensity[:, sel].sum(axis=1)
        if nt[ip]:
            prof[:, ip] *= 4.0 / nt[ip]
    return prof, nt, float(np.sum(tvals) / tvals.size)


def average_profiles(profiles: np.ndarray, nt: np.ndarray):
    """(average over the realizations with nt != 0, nsample) (src/main.f90:466-483)."""
    keep = np.asarray(nt) != 0
    nsample = int(keep.sum())
    avg = profiles[keep].sum(axis=0) / nsample if nsample else np.full(profiles.shape[1:], np.nan)
    return avg, nsample


def _pass(wguide, seed0: int, nseed: int, filtered: bool):
    plan = wguide.plan
    tv, it, st, code = plan.tprofile(None, seed0=seed0, nseed=nseed, filtered=filtered)
    if code:
        raise RecurGreenError(code)
    return tv, it


def compute_tprofile_batch(wguide, seed0: int, nseed: int, tmin: float, tmax: float):
    """(profiles[nseed, nx], nt[nseed]) of realizations seed0 .. seed0+nseed-1 (the OpenMP loop of
    task_tprofile, src/main.f90:440-461, as device passes)."""
    check_window(tmin, tmax)
    tv, it = _pass(wguide, seed0, nseed, False)
    out = np.zeros((nseed, wguide.plan.nx))
    nt = np.zeros(nseed, dtype=int)
    for r in range(nseed):
        out[r], nt[r] = window_profile(tv[r], it[r], tmin, tmax)
    return out, nt


def compute_tprofile(wguide, iseed: int, tmin: float, tmax: float):
    """compute_tprofile (src/waveguide.f90:843-907): (tprofile[nx], nt) of realization iseed."""
    p, nt = compute_tprofile_batch(wguide, iseed, 1, tmin, tmax)
    return p[0], int(nt[0])


def compute_itx_profile_batch(wguide, seed0: int, nseed: int, tmin: float, tmax: float):
    """itx[nseed, nx] of realizations seed0 .. seed0+nseed-1."""
    check_window(tmin, tmax)
    tv, it = _pass(wguide, seed0, nseed, False)
    wol = float(wguide.params["wol"])
    return np.stack([itx_window(tv[r], it[r], tmin, tmax, wol) for r in range(nseed)])


def compute_itx_profile(wguide, iseed: int, tmin: float, tmax: float):
    """compute_itx_profile (src/waveguide.f90:764-825): itx[nx] of realization iseed."""
    return compute_itx_profile_batch(wguide, iseed, 1, tmin, tmax)[0]


def compute_tprofile_group_batch(wguide, seed0: int, nseed: int, tm, dt):
    """(tprofile[nseed, nx, nprofile], nt[nseed, nprofile], tavg[nseed])."""
    check_group(tm, dt)
    tv, it = _pass(wguide, seed0, nseed, True)
    npf = np.asarray(tm).size
    out = np.zeros((nseed, wguide.plan.nx, npf))
    nt = np.zeros((nseed, npf), dtype=int)
    tavg = np.zeros(nseed)
    for r in range(nseed):
        out[r], nt[r], tavg[r] = group_profiles(tv[r], it[r], tm, dt)
    return out, nt, tavg


def compute_tprofile_group(wguide, iseed: int, tm, dt):
    """compute_tprofile_group (src/waveguide.f90:924-997): (tprofile[nx, nprofile], nt, tavg)."""
    p, nt, tavg = compute_tprofile_group_batch(wguide, iseed, 1, tm, dt)
    return p[0], nt[0], float(tavg[0])


def _g06(x: float) -> str:
    """Fortran g0.6 for the window bounds of the file headers (values in [0, 1])."""
    if x == 0.0:
        return "0.000000"
    if abs(x) < 0.1:
        m, e = f"{x:.5E}".split("E")
        mant = float(m) / 10.0
        return f"{mant:.6f}E{int(e) + 1:+03d}"
    digits_before = len(str(int(abs(x)))) if abs(x) >= 1 else 0
    return f"{x:.{6 - digits_before}f}"


def save_tprofile(directory: str, wguide, profiles: np.ndarray, avg: np.ndarray, nsample: int, tmin: float,
                  tmax: float, nseed: int, elapsed: float):
    """result_N.raw.csv (one profile per realization) and result_N.avg.csv ('x, tprofile') as
    save_tprofile (src/main.f90:500-545) writes them, without the TikZ step; returns both paths."""
    os.makedirs(directory, exist_ok=True)
    now = datetime.datetime.now().astimezone()
    head = ["%% Computed on " + now.strftime("%Y-%m-%d at %H:%M:%S %z") + " by RecurGreen-HIP"]
    head += parameter_lines(wguide, "%%")
    nx = profiles.shape[1]
    head.append(f"%% Transmission profile parameters: Found nsample={nsample} / nseed={nseed}, nx={nx}, "
                f"tmin={_g06(tmin)}, tmax={_g06(tmax)}, time={elapsed:.3f} s.")
    raw = unique_path(directory, "result_", ".raw.csv")
    with open(raw, "w") as f:
        f.write("\n".join(head) + "\n")
        f.write("%% Here follows, on each line, the function |psi_T(x)|^2 for a given realization of the disorder.\n")
        for row in profiles:
            f.write(", ".join(format_g016(v) for v in row) + "\n")
    avgp = unique_path(directory, "result_", ".avg.csv")
    xmesh = wguide.get_xmesh()
    with open(avgp, "w") as f:
        f.write("\n".join(head) + "\n")
        f.write("x, tprofile\n")
        for x, v in zip(xmesh, avg):
            f.write(f"{format_g016(x)}, {format_g016(v)}\n")
    return raw, avgp


def task_tprofile(wguide, nseed: int, tmin: float, tmax: float, outputdir: str | None = None, log=print):
    """task_tprofile (src/main.f90:402-493): profiles of realizations 1..nseed, their average over
    the realizations with nt > 0, and the two files under <outputdir>/tprofile/ (by default the
    reference's parameter-named folder, wguide.generate_directory()).  Returns (raw path, avg path,
    avg, nsample)."""
    if nseed <= 0:
        raise RecurGreenError(binding.RG_EINVAL, f"Invalid number of disorder realizations (received {nseed}), aborting...")
    t0 = time.perf_counter()
    profiles, nt = compute_tprofile_batch(wguide, 1, nseed, tmin, tmax)
    elapsed = time.perf_counter() - t0
    avg, nsample = average_profiles(profiles, nt)
    if nsample == 0:
        log(f"[WARN] Found no transmission eigenfunction (nsample={nsample}), results will be undetermined...")
    else:
        log(f"[INFO] Found {nsample}/{nseed} transmission eigenfunctions, now averaging...")
    if outputdir is None:
        outputdir = wguide.generate_directory()
    raw, avgp = save_tprofile(os.path.join(outputdir, "tprofile"), wguide, profiles, avg, nsample, tmin, tmax,
                              nseed, elapsed)
    return raw, avgp, avg, nsample
