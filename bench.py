"""Benchmark of the recursive Green function sweep on one MI355X.

Workload (BASELINE metric at N=128): Dirichlet waveguide, wavelength 1, W=64.25 (N=128 open
modes), L=50 with dx=dy=0.05 (nx=1000 slices, ny=1285 transverse cells), white-noise disorder
with l_s=10 (L/l_s=5), epsilon=0, a batch of 4096 realizations per step.  The disorder is
generated on the device from the seeds, so every input is resident in HBM when the timed
region starts; a step is one batched sweep (noise, projection, 1000 slice inversions and
Dyson GEMMs, Fisher-Lee, singular values) of the whole batch.

Prints one JSON line (rank 0).  `roofline` is measured live from HIP events recorded around
every launch of the dominant kernel on the plan's stream during the timed steps;
`cpu_baseline` times the FP64 CPU oracle (OpenMP, one realization per thread) on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "mixed_fp32": 157.3, "mixed_bf16x3": 2500.0}
DTYPE = {"fp64": "f64", "mixed_fp32": "f32+f64", "mixed_bf16x3": "bf16x3+f64"}


def workload_params(args):
    W, L, dx = 64.25, 50.0, 0.05
    nxdiso = int(round(L / dx)) if args.nx is None else args.nx
    return dict(boundtype="dirichlet", nxdiso=nxdiso, nxfree=0, wol=W, exfac=1.0, aratio=L / W,
                dscat=L / 10.0, dabso=0.0, napera=1.0, naperb=1.0, ny=int(round(W / 0.05)), dos_model=1)


def cpu_baseline(params, seconds):
    from oracle import oracle as O

    ow = O.Waveguide(**params)
    threads = O.omp_max_threads()
    # one realization per thread (schedule(dynamic,1) as src/main.f90:153); size the sample so it
    # takes about `seconds` of wall time given a one-realization probe on a reduced slice count
    probe_nx = max(20, params["nxdiso"] // 50)
    pp = dict(params, nxdiso=probe_nx, aratio=params["aratio"] * probe_nx / params["nxdiso"])
    t0 = O.wtime()
    O.Waveguide(**pp).tdistrib_batch(1, 1, nthreads=1)
    per_real = (O.wtime() - t0) * params["nxdiso"] / probe_nx
    nreal = max(threads, int(round(seconds / max(per_real, 1e-6))) * threads)
    nreal = min(nreal, 4 * threads)
    t0 = O.wtime()
    ow.tdistrib_batch(1, nreal, nthreads=threads)
    dt = O.wtime() - t0
    return {"value": nreal / dt, "unit": "realizations/s", "cores": threads, "kind": "port",
            "sample": f"{nreal} realizations of the same N=128, nx={params['nxdiso']}, ny={params['ny']} workload "
                      f"(FP64 oracle, OpenMP over realizations, {dt:.1f} s wall)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--nx", type=int, default=None, help="override slice count (profiling runs)")
    ap.add_argument("--precision", default="fp64", choices=list(PEAK_TFLOPS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the sweep shards by realization; see DESIGN.md); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch

    from recurgreen_amd import binding

    params = workload_params(args)
    plan = binding.Plan(params, batch_max=args.batch, device=0)
    n, nx, B = plan.nmu, plan.nx, args.batch
    out = torch.empty(B * plan.neigv, dtype=torch.float64, device="cuda:0")
    status = torch.zeros(B, dtype=torch.int32, device="cuda:0")

    for w in range(args.warmup):
        plan.tdistrib_device(1 + w * B, B, args.precision, out, status)
    plan.synchronize()
    torch.cuda.synchronize()

    plan.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        plan.tdistrib_device(1 + (args.warmup + s) * B, B, args.precision, out, status)
    plan.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    times = plan.kernel_times()
    plan.profile(False)
    bad = int((status != 0).sum().item())

    flops_slice = 16.0 * n ** 3
    flops_real = flops_slice * (nx - 1) + 8.0 * n ** 3
    value = args.steps * B / elapsed
    # dominant kernel and its roofline: algorithmic flops per launch / mean launch duration
    dom = max(("slice_inverse", "dyson_gemm"), key=lambda k: times.get(k, {}).get("ms", 0.0))
    launches = max(1, times[dom]["launches"])
    avg_ms = times[dom]["ms"] / launches
    per_launch = 8.0 * n ** 3 * B
    achieved = per_launch / (avg_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.precision]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(f"{dom}:{args.precision}")

    line = {
        "metric": "realizations/s at N=128 modes, mixed vs FP64 sweep; max rho(T) bin error",
        "value": value, "unit": "realizations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": DTYPE[args.precision], "data": "synthetic",
        "config": {"workload": f"dirichlet W=64.25 N={n} L=50 nx={nx} ny={plan.ny} ls=10 eps=0, "
                               f"batch {B} realizations, precision {args.precision}",
                   "global_batch": B, "parallelism": "realizations"},
        "slice_updates_per_s": value * (nx - 1),
        "algorithmic_tflops": value * flops_real / 1e12,
        "failed_realizations": bad,
        "kernel_ms_per_step": {k: v["ms"] / args.steps for k, v in times.items() if v["launches"]},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "per_launch_flops": per_launch, "avg_launch_ms": avg_ms},
    }
    plan.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(params, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
