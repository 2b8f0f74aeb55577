"""Autotuner of the sweep's launch options (no counterpart in the reference).

For one plan shape (boundary, mode count, transverse cells, batch, precision) it times short
sweeps of the real batch on the device -- the same N, ny and batch, the guide shortened to
`probe_slices` slices at the same dx and disorder strength -- under each candidate value of one
option at a time, keeps the fastest, and moves to the next option:

  inversion kernel   fp64: register-resident / look-ahead Gauss-Jordan / blocked LU;
                     mixed: verified diagonal pivots / look-ahead / blocked;
  Dyson kernel       mixed_fp32 with NP >= 80: one workgroup per realization / 64 x 64 tiles;
  projection split   1, 2, 4 or 8 partial sums over the transverse cells.

Every candidate computes the same quantity (the kernels differ in tiling and pivoting only), so
the choice changes speed, not the meaning of the result.  Results are cached in a JSON file keyed
by the device name and the shape; `apply` sets a plan's options from the cache.  The two-lane
batch split is left to the caller (it overlaps kernels of two streams, which makes per-kernel
timings overlap too).
"""
from __future__ import annotations

import json
import os
import time

from .binding import Plan

CACHE = os.environ.get("RG_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_cache.json"))

CANDIDATES = {
    "fp64": [("inverse_fp64", (4, 3, 0, 1, 2)), ("dyson_fp64", (1, 2, 0)), ("project_parts", (1, 2, 4, 8))],
    "mixed_fp32": [("inverse_mixed", (4, 3, 1, 2)), ("dyson", (1, 3, 0)), ("project_parts", (1, 2, 4, 8))],
    "mixed_bf16x3": [("inverse_mixed", (4, 3, 1, 2)), ("project_parts", (1, 2, 4, 8))],
}


def device_name(device: int = 0) -> str:
    import torch

    return torch.cuda.get_device_name(device)


def shape_key(plan: Plan, params: dict, batch: int, precision: str, device: int = 0) -> str:
    return (f"{device_name(device)}|{params.get('boundtype', 'dirichlet')}|nmu={plan.nmu}|ny={plan.ny}"
            f"|batch={batch}|{precision}")


def probe_params(params: dict, probe_slices: int) -> dict:
    """The same waveguide cut to `probe_slices` disordered slices: aratio and dscat scale with the
    slice count, so kdx and the per-cell disorder strength do not change."""
    f = probe_slices / float(params["nxdiso"])
    return dict(params, nxdiso=probe_slices, aratio=params.get("aratio", 1.0) * f,
                dscat=params.get("dscat", 0.0) * f)


def _time_sweep(plan: Plan, precision: str, batch: int, reps: int) -> float:
    import torch

    out = torch.empty(batch * plan.neigv, dtype=torch.float64, device="cuda")
    st = torch.zeros(batch, dtype=torch.int32, device="cuda")
    plan.tdistrib_device(1, batch, precision, out, st)  # warm-up (code objects, caches)
    plan.synchronize()
    best = float("inf")
    for r in range(reps):
        t0 = time.perf_counter()
        plan.tdistrib_device(1 + r * batch, batch, precision, out, st)
        plan.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def load_cache(path: str = CACHE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def tune(params: dict, batch: int, precision: str, probe_slices: int = 20, reps: int = 2, device: int = 0,
         path: str | None = CACHE, log=None) -> dict:
    """Time the candidates on a probe plan and return {option: value, "times": {...}}; the entry
    is stored in the cache at `path` (None: not stored)."""
    probe = Plan(probe_params(params, probe_slices), batch_max=batch, device=device)
    try:
        chosen, times = {}, {}
        for key, values in CANDIDATES[precision]:
            if key in ("dyson", "dyson_fp64") and probe.nmu and (probe.nmu + 15) // 16 * 16 < 80:
                continue  # only the 64 x 64 kernel exists below NP = 80
            res = {}
            for v in values:
                probe.set_option(key, v)
                res[v] = _time_sweep(probe, precision, batch, reps)
                if log:
                    log(f"[autotune] {precision} {key}={v}: {res[v] * 1e3:.2f} ms for {probe_slices} slices")
            best = min(res, key=res.get)
            probe.set_option(key, best)
            chosen[key] = best
            times[key] = {str(v): t for v, t in res.items()}
        entry = dict(chosen, times=times, probe_slices=probe_slices)
        if path:
            cache = load_cache(path)
            cache[shape_key(probe, params, batch, precision, device)] = entry
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(cache, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        return entry
    finally:
        probe.close()


def apply(plan: Plan, params: dict, batch: int, precision: str, device: int = 0, path: str = CACHE) -> dict | None:
    """Set the plan's options from the cache entry of its shape; returns the entry or None."""
    entry = load_cache(path).get(shape_key(plan, params, batch, precision, device))
    if not entry:
        return None
    for key, _ in CANDIDATES[precision]:
        if key in entry:
            plan.set_option(key, entry[key])
    return entry
