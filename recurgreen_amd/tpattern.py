"""Radiation pattern of transmitted and reflected probabilities per channel (task `tpattern`).

Mirrors task_tpattern and save_tpattern of src/main.f90:735-901 for probatype 'incident'
(|t(:, imu)|^2 and |r(:, imu)|^2 for the fixed incident channel imu) and 'total' (column sums
sum_b |t_ba|^2 and sum_b |r_ba|^2), averaged over realizations 1..nseed.  The reference computes
one compute_refl_trans_v1 per seed inside an OpenMP loop (:790-835); here the seeds run as
batched complex128 device passes of the plan's batch size, and only the per-channel
probabilities come back to the host.  probatype 'deposition' (:808-812) takes |dep(:, imu)|^2 of
the deposition matrix between the input lead (ipos = 1) and the middle slice fpos = nxfree +
nxdiso/2 + 1 (:765), from rg_deposition; its reflection column is zero.
"""
from __future__ import annotations

import datetime
import os
import time

import numpy as np

from . import binding
from .binding import RecurGreenError
from .distrib import _parse_fortran_real, format_g016

PROBATYPES = ("incident", "total", "deposition")


def mode_index(boundtype: str, nmu: int, imode: int) -> int:
    """Channel index imu (1-based) of incident mode `imode` (src/main.f90:772-776): in the
    periodic case the channels are ordered by transverse momentum, so imode=1 (the frontal mode)
    sits at the middle of the list."""
    if boundtype == "periodic":
        return (imode + nmu // 2 - 1) % nmu + 1
    return imode


def check_arguments(nseed: int, probatype: str) -> None:
    """Argument checks of src/main.f90:754-761."""
    if nseed <= 0:
        raise RecurGreenError(binding.RG_EINVAL,
                              f"Invalid number of disorder realizations (received {nseed}), aborting...")
    if probatype not in PROBATYPES:
        raise RecurGreenError(binding.RG_EINVAL, f"Invalid probatype, received '{probatype}', "
                              "expected incident/total/deposition, aborting...")


def channel_probabilities(rmat: np.ndarray, tmat: np.ndarray, probatype: str, imu: int):
    """Per-realization (tprob, rprob), each (nreal, nmu), from r and t as (nreal, nmu, nmu) with
    [r, b, a] = matrix element (b, a) (src/main.f90:815-821)."""
    if probatype == "incident":
        return np.abs(tmat[:, :, imu - 1]) ** 2, np.abs(rmat[:, :, imu - 1]) ** 2
    return np.sum(np.abs(tmat) ** 2, axis=1), np.sum(np.abs(rmat) ** 2, axis=1)


def deposition_positions(params: dict) -> tuple[int, int]:
    """(fpos, ipos) of the deposition pattern (src/main.f90:769-770): the middle of the medium, from
    the first slice."""
    return int(params.get("nxfree", 0)) + int(params["nxdiso"]) // 2 + 1, 1


def compute_tpattern(wguide, nseed: int, imode: int, probatype: str = "incident", seed0: int = 1):
    """(tprob, rprob, imu): probabilities per outgoing channel averaged over realizations
    seed0 .. seed0+nseed-1 (src/main.f90:802-845).  `wguide` is a recurgreen_amd.Waveguide."""
    check_arguments(nseed, probatype)
    plan = wguide.plan
    nmu = plan.nmu
    imu = mode_index(wguide.params["boundtype"], nmu, imode)
    if probatype in ("incident", "deposition") and not 1 <= imu <= nmu:
        raise RecurGreenError(binding.RG_EINVAL, f"Invalid incident mode imode={imode}, expected in [1, {nmu}]")
    tsum = np.zeros(nmu)
    rsum = np.zeros(nmu)
    chunk = max(1, plan.batch_max)
    fpos, ipos = deposition_positions(wguide.params)
    for off in range(0, nseed, chunk):
        nb = min(chunk, nseed - off)
        if probatype == "deposition":
            d, st, code = plan.deposition(fpos, ipos, None, seed0=seed0 + off, nseed=nb)
            if code:
                raise RecurGreenError(code)
            tsum += (np.abs(d[:, :, imu - 1]) ** 2).sum(axis=0)
            continue
        r, t, st, code = plan.refl_trans(None, seed0=seed0 + off, nseed=nb)
        if code:
            raise RecurGreenError(code)
        tp, rp = channel_probabilities(r, t, probatype, imu)
        tsum += tp.sum(axis=0)
        rsum += rp.sum(axis=0)
    return tsum / nseed, rsum / nseed, imu


def parameter_lines(wguide, prefix: str = "%%") -> list[str]:
    """The three lines of print_parameters_waveguide (src/waveguide.f90:1003-1018)."""
    p, pl = wguide.params, wguide.plan

    def f6(x):
        return f"{x:.6f}"

    return [
        f"{prefix} Waveguide 2D with boundtype='{p['boundtype']}', W/lambda={f6(p['wol'])}, exfac={f6(p['exfac'])}, "
        f"Nmode={pl.nmu}, Nprop={pl.nprop}",
        f"{prefix} Filtering: NAper_a={f6(p['napera'])} (Nmode_a={pl.jmax - pl.jmin + 1}), "
        f"NAper_b={f6(p['naperb'])} (Nmode_b={pl.imax - pl.imin + 1}), Neigv={pl.neigv}",
        f"{prefix} Medium: L/lscat={f6(p['dscat'])}, L/labso={f6(p['dabso'])}, Nx_diso={p['nxdiso']}, "
        f"Nx_free={p['nxfree']}, aratio=L/W={f6(p['aratio'])}",
    ]


def save_tpattern(path: str, wguide, tprob, rprob, probatype: str, imode: int, nseed: int, elapsed: float) -> None:
    """CSV of save_tpattern (src/main.f90:882-893): %% header, 'imu, tprob, rprob', one row per
    channel with g0.16 values.  The TikZ step (:895-899) is not part of the library."""
    now = datetime.datetime.now().astimezone()
    lines = ["%% Computed on " + now.strftime("%Y-%m-%d at %H:%M:%S %z") + " by RecurGreen-HIP"]  # print_timestamp
    lines += parameter_lines(wguide, "%%")
    lines.append(f"%% Radiation pattern parameters: probatype='{probatype}', nseed={nseed}, imode={imode}, "
                 f"time={elapsed:.3f} s.")
    lines.append("imu, tprob, rprob")
    lines += [f"{i + 1}, {format_g016(tprob[i])}, {format_g016(rprob[i])}" for i in range(len(tprob))]
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def load_tpattern(path: str) -> np.ndarray:
    """(nmu, 3) array [imu, tprob, rprob] from a tpattern CSV (header lines skipped)."""
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("%") or s.startswith("imu"):
                continue
            rows.append([_parse_fortran_real(x.strip()) for x in s.split(",")])
    return np.array(rows)


def task_tpattern(wguide, nseed: int, imode: int, probatype: str = "incident", outputdir: str | None = None):
    """Compute and save under <outputdir>/tpattern/ (by default the reference's parameter-named
    folder, wguide.generate_directory()): returns (path, tprob, rprob)."""
    t0 = time.perf_counter()
    tprob, rprob, _ = compute_tpattern(wguide, nseed, imode, probatype)
    elapsed = time.perf_counter() - t0
    if outputdir is None:
        outputdir = wguide.generate_directory()
    d = os.path.join(outputdir, "tpattern")
    k = 1
    while os.path.exists(os.path.join(d, f"result_{k}.csv")):  # unique_filename (src/base_utils.f90:825-840)
        k += 1
    path = os.path.join(d, f"result_{k}.csv")
    save_tpattern(path, wguide, tprob, rprob, probatype, imode, nseed, elapsed)
    return path, tprob, rprob
