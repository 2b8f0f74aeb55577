"""Benchmark of the recursive Green function sweep on one MI355X.

Workload (the headline metric, at N=128): Dirichlet waveguide, wavelength 1, W=64.25 (N=128 open
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

METRIC = "realizations/s at N=128 modes, mixed vs FP64 sweep; max rho(T) bin error"
PEAK_TFLOPS = {"fp64": 78.6, "mixed_fp32": 157.3, "mixed_bf16x3": 2500.0}
DTYPE = {"fp64": "f64", "mixed_fp32": "f32+f64", "mixed_bf16x3": "bf16x3+f64"}


# Waveguides of the benchmark: width W, slices nx (L = nx dx), l_s, batch.  n128 is the headline
# workload; c1..c4 are the other configurations of the task (c1 CPU-sized, c2 FP64 N=100, c3 mixed
# N=256, c4 the adjoint's N=64).
CONFIGS = {
    "n128": (64.25, 1000, 10.0, 4096),
    "c1": (10.25, 200, 5.0, 1000),
    "c2": (50.25, 1000, 10.0, 8192),
    "c3": (128.25, 2000, 20.0, 4096),
    "c4": (32.25, 400, 10.0, 256),
}


def workload_params(args):
    """Dirichlet guide of the chosen configuration at dx=dy=0.05 (L = nx dx, dscat = L / l_s);
    --nx shortens it at fixed dx."""
    W, nx0, ls, _ = CONFIGS[getattr(args, "config", "n128")]
    dx = 0.05
    nxdiso = nx0 if args.nx is None else args.nx
    L = nxdiso * dx
    return dict(boundtype="dirichlet", nxdiso=nxdiso, nxfree=0, wol=W, exfac=1.0, aratio=L / W,
                dscat=L / ls, dabso=0.0, napera=1.0, naperb=1.0, ny=int(round(W / dx)), dos_model=1)


def kernel_flops(plan, params, kernel, prec) -> float:
    """Algorithmic real flops per realization of one launch of `kernel` at N modes: 8 N^3 for the Dyson
    product and the general inversion (zgetrf + zgetri); for the symmetric inversion of Dirichlet
    guides the lower-triangle tiles of 16: 4 N^2 (N + 16) (sweep-operator updates, panel products and the
    diagonal-block inversions), directly up to N = 128 and for complex64 sweeps up to 256; complex128
    sweeps above 128 modes invert through the Schur complement, whose symmetric form (blocks n1 = 128
    and n2 = N - 128) does two symmetric block inversions and four products, 16 n1 n2 (n1 + n2)."""
    n = plan.nmu
    npad = 16 * ((n + 15) // 16)  # the kernels' padded order
    if kernel == "slice_inverse" and params.get("boundtype", "dirichlet") == "dirichlet" and 32 <= npad:
        if plan.get_option("inverse_fp64" if prec == "fp64" else "inverse_mixed") == 4:
            if npad <= 128 or prec != "fp64":
                return 4.0 * n * n * (n + 16)
            n1, n2 = 128, n - 128
            if n2 >= 32:
                return 4.0 * n1 * n1 * (n1 + 16) + 4.0 * n2 * n2 * (n2 + 16) + 16.0 * n1 * n2 * (n1 + n2)
    return 8.0 * n ** 3


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, seconds):
    """The reference's FP64 slice loop on the host's LAPACK/BLAS (oracle/lapack_sweep.py: zgetrf + zgetri
    and zgemm per slice, batched over realizations on MKL through torch's CPU backend), timed on a bounded
    sample of the same workload; the textbook C restatement (oracle/rg_oracle.c, OpenMP over realizations)
    is timed beside it on a smaller sample."""
    from oracle import oracle as O
    from oracle.lapack_sweep import tdistrib_lapack

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    nx = params["nxdiso"]
    # size the sample from a short guide of the same width (time per realization scales with nx)
    probe_nx = max(20, nx // 20)
    pp = dict(params, nxdiso=probe_nx, aratio=params["aratio"] * probe_nx / nx, dscat=params["dscat"] * probe_nx / nx)
    t0 = time.perf_counter()
    tdistrib_lapack(O.Waveguide(**pp), range(1, 2 * threads + 1), threads)
    per_real = (time.perf_counter() - t0) * nx / probe_nx / (2 * threads)
    nreal = int(min(8 * threads, max(2 * threads, seconds / max(per_real, 1e-9))))
    ow = O.Waveguide(**params)
    t0 = time.perf_counter()
    tdistrib_lapack(ow, range(1, nreal + 1), threads)
    dt = time.perf_counter() - t0
    # the C restatement: one realization per thread, a quarter of the sample
    nc = max(threads, nreal // 4 // threads * threads)
    t1 = O.wtime()
    ow.tdistrib_batch(1, nc, nthreads=threads)
    dtc = O.wtime() - t1
    return {"value": nreal / dt, "unit": "realizations/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{nreal} realizations of the same N={ow.nmu}, nx={nx}, ny={params['ny']} workload: the "
                      f"reference's slice loop on LAPACK/BLAS (MKL zgetrf+zgetri, zgemm, gesdd through torch's CPU "
                      f"backend, {threads} threads), {dt:.1f} s wall",
            "textbook_c_port": {"value": nc / dtc, "sample": f"{nc} realizations, oracle/rg_oracle.c, OpenMP over "
                                                             f"realizations, {dtc:.1f} s wall"}}


def timed_sweeps(plan, precision, B, steps, warmup, out, status):
    """Warmup + timed batched sweeps; the same seeds for every precision.  The warmup steps record
    every kernel (the per-kernel breakdown, per step) and name the dominant kernel; the timed steps
    record HIP events around the dominant kernel's launches only, so that the events of the other
    launches do not slow the timed region.  Returns (elapsed s, dominant kernel's times over the
    timed region, breakdown, dominant name, failed realizations); `out` holds the last step's
    eigenvalues."""
    import torch

    plan.set_option("profile_mask", -1)
    plan.profile(warmup > 0)
    for w in range(warmup):
        plan.tdistrib_device(1 + (w % max(1, steps)) * B, B, precision, out, status)
    plan.synchronize()
    torch.cuda.synchronize()
    warm = plan.kernel_times() if warmup > 0 else {}
    plan.profile(False)
    if warm:
        dom = max(("slice_inverse", "dyson_gemm"), key=lambda k: warm.get(k, {}).get("ms", 0.0))
        breakdown = {k: v["ms"] / warmup for k, v in warm.items() if v["launches"]}
        plan.set_option("profile_mask", 1 << list(warm).index(dom))
    else:
        dom, breakdown = None, None
    plan.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        plan.tdistrib_device(1 + s * B, B, precision, out, status)
    plan.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    times = plan.kernel_times()
    plan.profile(False)
    plan.set_option("profile_mask", -1)
    if dom is None:  # no warmup: the timed steps recorded every kernel
        dom = max(("slice_inverse", "dyson_gemm"), key=lambda k: times.get(k, {}).get("ms", 0.0))
        breakdown = {k: v["ms"] / steps for k, v in times.items() if v["launches"]}
    return elapsed, times[dom], breakdown, dom, int((status != 0).sum().item())


def accuracy(plan, t_ref, t_mix):
    """Errors of a mixed sweep against the FP64 sweep on identical seeds: mean T, max |dT|, relative errors
    by decade of T, and the max bin deviation of rho(T) on 50 linear bins and on the Chebyshev bins of the
    tdistrib task (50 and 256), with the largest count difference of a bin and what one count changes in
    the narrowest bin (sum of counts times the bin width)."""
    from recurgreen_amd import distrib

    out = {"mean_T_fp64": float(t_ref.mean().item()), "mean_T_err": float((t_mix.mean() - t_ref.mean()).abs().item()),
           "mean_T_shift": float((t_mix.mean() - t_ref.mean()).item()),
           "max_abs_dT": float((t_mix - t_ref).abs().max().item())}
    out["fp64_T_above_1"] = int((t_ref > 1.0).sum().item())
    out["mixed_T_at_1"] = int((t_mix == 1.0).sum().item())
    rel = (t_mix - t_ref).abs() / t_ref.clamp_min(1e-300)
    for lo, hi in ((1e-3, 1.01), (1e-6, 1e-3), (1e-9, 1e-6)):
        sel = (t_ref > lo) & (t_ref <= hi)
        if int(sel.sum().item()):
            out[f"rel_err_rms_T_{lo:g}_{hi:g}"] = float(rel[sel].pow(2).mean().sqrt().item())
    for bt, nb in (("linear", 50), ("chebyshev", 50), ("chebyshev", 256)):
        e = distrib.bins(bt, nb)
        cr = plan.histogram(t_ref, e)[0]
        cm = plan.histogram(t_mix, e)[0]
        rr = distrib.rho_from_counts(cr, e)[:, 1]
        rm = distrib.rho_from_counts(cm, e)[:, 1]
        d = abs(rm - rr)
        out[f"max_rho_bin_err_{bt}{nb}"] = float(d.max())
        out[f"max_rho_bin_{bt}{nb}"] = int(d.argmax())
        out[f"max_count_diff_{bt}{nb}"] = int(abs(cm.astype("int64") - cr.astype("int64")).max())
        out[f"rho_of_one_count_narrowest_{bt}{nb}"] = float(1.0 / (cr.sum() * (e[1:] - e[:-1]).min()))
    return out


def uneven_workload(nreal, seed=12345):
    """Uneven batch: N ~ U{16,32,64,128,256}, W = N/2 + 1/8, L/l_s ~ U[0.5, 10] with l_s = 10,
    eps ~ U[0, 0.05 k^2] (dabso = eps L / k), dx = dy = 0.05, white-noise disorder; seeds 1..nreal."""
    import numpy as np

    rng = np.random.default_rng(seed)
    k, dx, ls = 2 * np.pi, 0.05, 10.0
    out = []
    for _ in range(nreal):
        N = int(rng.choice([16, 32, 64, 128, 256]))
        W = N / 2 + 0.125
        L = ls * rng.uniform(0.5, 10.0)
        nx = max(2, int(round(L / dx)))
        L = nx * dx
        eps = rng.uniform(0.0, 0.05 * k * k)
        out.append(dict(boundtype="dirichlet", nxdiso=nx, nxfree=0, wol=W, exfac=1.0, aratio=L / W, dscat=L / ls,
                        dabso=eps * L / k, napera=1.0, naperb=1.0, ny=int(round(W / dx)), dos_model=1))
    return out


def uneven_bench(args):
    """One rg_tdistrib_uneven call per precision over the same uneven batch (wall clock around the
    call: host set-up, one plan per width, the sweeps and the copies back)."""
    import numpy as np
    import torch

    from recurgreen_amd import binding

    torch.cuda.init()
    ps = uneven_workload(args.uneven)
    seeds = np.arange(1, len(ps) + 1, dtype=np.int64)
    Ns = np.array([int(np.floor(2 * p["wol"])) for p in ps])
    res = {}
    for prec in ("fp64", args.precision):
        t0 = time.perf_counter()
        vals, st, code = binding.tdistrib_uneven(ps, seeds, prec, batch_max=args.batch)
        dt = time.perf_counter() - t0
        res[prec] = (vals, st, dt)
        print(f"[bench-uneven] {prec}: {dt:.1f} s", file=sys.stderr, flush=True)
    v64, s64, t64 = res["fp64"]
    vm, sm, tm = res[args.precision]
    ok = (s64 == 0) & (sm == 0)
    flops = sum(16.0 * int(Ns[r]) ** 3 * (p["nxdiso"] - 1) + 8.0 * int(Ns[r]) ** 3 for r, p in enumerate(ps) if ok[r])
    per_n = {}
    for N in sorted(set(Ns.tolist())):
        idx = [r for r in range(len(ps)) if Ns[r] == N]
        good = [r for r in idx if ok[r]]
        err = max((float(np.max(np.abs(vm[r] - v64[r]))) for r in good), default=None)
        per_n[str(N)] = {"realizations": len(idx), "swept": len(good), "max_abs_dT_mixed_vs_fp64": err,
                         "status": sorted(set(int(s64[r]) for r in idx))}
    line = {
        "metric": "realizations/s, uneven batch (c5: N in {16..256}, L/ls in [0.5, 10], eps in [0, 0.05 k^2])",
        "value": int(ok.sum()) / tm, "unit": "realizations/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": DTYPE[args.precision], "data": "synthetic",
        "config": {"workload": f"uneven batch of {len(ps)} realizations, seeds 1..{len(ps)}, batch_max {args.batch}, "
                               f"precision {args.precision}"},
        "fp64_value": int(ok.sum()) / t64, "speedup_vs_fp64": t64 / tm,
        "algorithmic_tflops": flops / tm / 1e12, "fp64_algorithmic_tflops": flops / t64 / 1e12,
        "per_mode_count": per_n,
    }
    print(json.dumps(line), flush=True)


def adjoint_bench(args):
    """Gradient of L = Tr(t^H t) and of T_max for the c4 waveguide (N=64, 400 slices, ny=645, batch
    256 by default): one adjoint call per loss and precision on device potentials, timed after a
    warmup call; the complex64 gradient is compared with the complex128 one."""
    import numpy as np
    import torch

    from recurgreen_amd import binding

    params = workload_params(args)
    B = args.batch
    plan = binding.Plan(params, batch_max=B)
    plan.adjoint_reserve(B)
    U = plan.generate_potential(1, B)
    u = torch.from_numpy(np.ascontiguousarray(U.transpose(2, 1, 0))).cuda()
    out = {}
    for loss in ("g", "tmax"):
        res = {}
        for prec in ("fp64", "mixed_fp32"):
            plan.set_option("adjoint_precision", binding.PRECISIONS[prec])
            plan.adjoint_device(u, loss)
            plan.synchronize()
            t0 = time.perf_counter()
            val, grad = plan.adjoint_device(u, loss)
            plan.synchronize()
            dt = time.perf_counter() - t0
            res[prec] = (dt, val.cpu().numpy(), grad.cpu().numpy())
        d64, v64, g64 = res["fp64"]
        d32, v32, g32 = res["mixed_fp32"]
        out[loss] = {"fp64_gradients_per_s": B / d64, "mixed_fp32_gradients_per_s": B / d32, "speedup": d64 / d32,
                     "grad_rel_err_fro": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
                     "loss_max_rel_err": float(np.max(np.abs(v32 - v64) / np.abs(v64)))}
    line = {"metric": "adjoint gradients/s (dL/dU over the whole grid), complex128 and complex64",
            "value": out["g"]["mixed_fp32_gradients_per_s"], "unit": "gradients/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": f"{args.config}: W={params['wol']} nx={params['nxdiso']} ny={params['ny']} batch {B}"},
            "losses": out}
    plan.close()
    print(json.dumps(line), flush=True)


def tprofile_bench(args):
    """Transmission eigenstates inside the medium (rg_tprofile, the data of compute_tprofile) for the
    c4 waveguide by default: one complex128 pass over a batch of seeds, timed after a warmup pass,
    host copies of T and of the per-slice intensities included."""
    from recurgreen_amd import binding

    params = workload_params(args)
    B = args.batch
    plan = binding.Plan(params, batch_max=B)
    plan.tprofile_reserve(B)
    plan.tprofile(None, seed0=1, nseed=B)
    plan.synchronize()
    t0 = time.perf_counter()
    tv, it, st, code = plan.tprofile(None, seed0=1, nseed=B)
    dt = time.perf_counter() - t0
    n, nx = plan.nmu, plan.nx
    flops = (32.0 * n ** 3 * nx) * B  # sweep (inversion + Dyson) and back substitution (two products) per slice
    line = {"metric": "tprofile realizations/s (eigenvalues of t^H t and |G(x;1) s U(:,v)|^2 at every slice), complex128",
            "value": B / dt, "unit": "realizations/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic", "failed_realizations": int((st != 0).sum()),
            "config": {"workload": f"{args.config}: W={params['wol']} N={n} nx={nx} ny={params['ny']} batch {B}"},
            "algorithmic_tflops": flops / dt / 1e12}
    plan.close()
    print(json.dumps(line), flush=True)


def refine_bench(args):
    """The complex64 sweep with `--refine K` FP64 refinement passes of G(L;0) (plan option refine) on the
    N=128 workload, beside the complex128 and plain complex64 sweeps of the same batch (192 realizations by
    default: the refinement keeps 3 nx N^2 complex128 per realization, 0.79 GB at nx = 1000, so passes hold
    about 200); each precision timed over one call after a warmup call, host copies included.  The refined
    eigenvalues are compared with the complex128 sweep's, with the rho(T) bin errors of accuracy() on the
    linear and Chebyshev bins."""
    import numpy as np

    from recurgreen_amd import binding

    params = workload_params(args)
    B = args.batch
    plan = binding.Plan(params, batch_max=B)
    res = {}
    for label, prec, passes in (("fp64", "fp64", 0), ("mixed_fp32", "mixed_fp32", 0),
                                (f"mixed_fp32_refine{args.refine}", "mixed_fp32", args.refine)):
        plan.set_option("refine", passes)
        plan.tdistrib(1, B, prec)
        t0 = time.perf_counter()
        T, st, code = plan.tdistrib(1, B, prec)
        dt = time.perf_counter() - t0
        res[label] = (B / dt, T.reshape(B, -1), int((st != 0).sum()))
    import torch

    t64 = torch.from_numpy(np.ascontiguousarray(res["fp64"][1]).ravel()).cuda()
    acc = {label: accuracy(plan, t64, torch.from_numpy(np.ascontiguousarray(v[1]).ravel()).cuda())
           for label, v in res.items() if label != "fp64"}
    plan.close()
    T64 = res["fp64"][1]
    err = {}
    for label in res:
        if label == "fp64":
            continue
        T = res[label][1]
        tmax = T64.max(axis=1, keepdims=True)
        ok = T64 >= 1e-11 * tmax
        bound = 1e-10 * T64 + 2e-13 * np.sqrt(T64 * tmax) + 1e-26 * tmax
        err[label] = {"max_rel_err_resolved": float(np.max(np.abs(T - T64)[ok] / T64[ok])),
                      "max_err_over_fp64_bound": float(np.max(np.abs(T - T64) / bound)),
                      "mean_T_shift": float(T.mean() - T64.mean())}
    label = f"mixed_fp32_refine{args.refine}"
    line = {"metric": "realizations/s of the complex64 sweep with FP64 refinement of G(L;0)",
            "value": res[label][0], "unit": "realizations/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": f"{args.config}: W={params['wol']} nx={params['nxdiso']} ny={params['ny']} batch {B}"},
            "realizations_per_s": {k: v[0] for k, v in res.items()},
            "failed_realizations": {k: v[2] for k, v in res.items()}, "against_fp64_sweep": err,
            "accuracy_against_fp64_sweep": acc}
    print(json.dumps(line), flush=True)


def copy_calibration(mib):
    """Plain device-to-device copies of `mib` MiB (torch's copy kernel): the reference kernel whose
    FETCH_SIZE / WRITE_SIZE counters scale the sweep kernels' counters to bytes."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.randn(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                      "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="n128", choices=list(CONFIGS),
                    help="waveguide configuration (default: the N=128 headline workload)")
    ap.add_argument("--batch", type=int, default=None, help="realizations per step (default: the configuration's)")
    ap.add_argument("--nx", type=int, default=None, help="override slice count (profiling runs)")
    ap.add_argument("--precision", default="mixed_fp32", choices=list(PEAK_TFLOPS),
                    help="sweep reported as `value`; the FP64 sweep is always measured beside it")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--uneven", type=int, default=0, metavar="K",
                    help="measure an uneven batch of K realizations (rg_tdistrib_uneven) instead")
    ap.add_argument("--refine", type=int, default=0, metavar="K",
                    help="time the complex64 sweep with K FP64 refinement passes (batch 192 by default)")
    ap.add_argument("--adjoint", action="store_true", help="time the adjoint (gradient) sweep instead (c4 by default)")
    ap.add_argument("--tprofile", action="store_true",
                    help="time the transmission-eigenstate pass (rg_tprofile) instead (c4 by default)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch options instead of the autotuner's (recurgreen_amd.autotune)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run plain copies of MIB MiB (traffic-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (realizations are independent and would shard by seed "
              f"range across GPUs); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if (args.adjoint or args.tprofile) and args.config == "n128":
        args.config = "c4"
    if args.batch is None:
        args.batch = 192 if args.refine else CONFIGS[args.config][3]
    if args.refine:
        refine_bench(args)
        return
    if args.adjoint:
        adjoint_bench(args)
        return
    if args.tprofile:
        tprofile_bench(args)
        return
    if args.uneven:
        uneven_bench(args)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import torch

    from recurgreen_amd import binding

    params = workload_params(args)
    plan = binding.Plan(params, batch_max=args.batch, device=0)
    n, nx, B = plan.nmu, plan.nx, args.batch
    flops_real = 16.0 * n ** 3 * (nx - 1) + 8.0 * n ** 3
    status = torch.zeros(B, dtype=torch.int32, device="cuda:0")
    runs = {}
    order = ["fp64"] + ([args.precision] if args.precision != "fp64" else [])
    outs = {}
    options = {}
    for prec in order:
        if not args.no_tune:
            from recurgreen_amd import autotune

            entry = autotune.apply(plan, params, B, prec) or autotune.tune(
                params, B, prec, log=lambda m: print(m, file=sys.stderr, flush=True))
            autotune.apply(plan, params, B, prec)
            options[prec] = {k: v for k, v in entry.items() if k in dict(autotune.CANDIDATES[prec])}
        outs[prec] = torch.empty(B * plan.neigv, dtype=torch.float64, device="cuda:0")
        el, dtimes, breakdown, dom, bad = timed_sweeps(plan, prec, B, args.steps, args.warmup, outs[prec], status)
        launches = max(1, dtimes["launches"])
        avg_ms = dtimes["ms"] / launches
        # realizations per launch: the batch, or half of it when the plan sweeps two lanes
        per_step = nx if dom == "slice_inverse" else nx - 1
        lanes = max(1, round(launches / (args.steps * per_step)))
        per_launch = kernel_flops(plan, params, dom, prec) * B / lanes
        runs[prec] = {"value": args.steps * B / el, "ms_per_step": el / args.steps * 1e3, "failed_realizations": bad,
                      "kernel_ms_per_step": breakdown,
                      "dominant": dom, "avg_launch_ms": avg_ms, "realizations_per_launch": B // lanes,
                      "achieved_tflops": per_launch / (avg_ms * 1e-3) / 1e12,
                      "per_realization_flops": kernel_flops(plan, params, dom, prec)}
        print(f"[bench] {prec}: {runs[prec]['value']:.1f} realizations/s, {runs[prec]['ms_per_step']:.0f} ms/step",
              file=sys.stderr, flush=True)
    main_prec = order[-1]
    r = runs[main_prec]
    # the same steps as captured HIP graphs (no events inside): one capture, then replays
    plan.set_option("graph", 1)
    plan.tdistrib_device(1, B, main_prec, outs[main_prec], status)
    plan.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        plan.tdistrib_device(1 + s * B, B, main_prec, outs[main_prec], status)
    plan.synchronize()
    torch.cuda.synchronize()
    gel = time.perf_counter() - t0
    plan.set_option("graph", 0)
    graph_run = {"value": args.steps * B / gel, "ms_per_step": gel / args.steps * 1e3,
                 "note": "same steps replayed from a captured graph, no events recorded"}
    peak = PEAK_TFLOPS[main_prec]
    traffic = traffic64 = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))  # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counters
        traffic = tj.get(f"{args.config}:{r['dominant']}:{main_prec}")
        traffic64 = tj.get(f"{args.config}:{runs['fp64']['dominant']}:fp64")
    line = {
        "metric": METRIC if args.config == "n128" else f"realizations/s at N={n} modes ({args.config})",
        "value": r["value"], "unit": "realizations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "data": "synthetic",
        # option dyson = 2: the complex64 Dyson product on BF16 matrix cores with three-way split FP32 operands
        "dtype": "f32+bf16x6+f64" if (options or {}).get(main_prec, {}).get("dyson") == 2 else DTYPE[main_prec],
        "config": {"workload": f"{args.config}: dirichlet W={params['wol']} N={n} L={nx * 0.05:g} nx={nx} ny={plan.ny} "
                               f"ls={CONFIGS[args.config][2]:g} eps=0, batch {B} realizations, precision {main_prec}",
                   "global_batch": B, "parallelism": "realizations"},
        "slice_updates_per_s": r["value"] * (nx - 1),
        # the reference algorithm's flops (general inversion and Dyson product, 16 N^3 per slice) per second
        "algorithmic_tflops": r["value"] * flops_real / 1e12,
        "fp64_sweep": runs["fp64"],
        "fp64_roofline": {"bound": "mfma", "kernel": runs["fp64"]["dominant"], "achieved": runs["fp64"]["achieved_tflops"],
                          "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                          "frac": runs["fp64"]["achieved_tflops"] / PEAK_TFLOPS["fp64"],
                          "traffic": traffic64, "avg_launch_ms": runs["fp64"]["avg_launch_ms"]},
        # dense matrix-core peaks of the two precisions: the ceiling of the mixed/FP64 ratio for sweeps whose
        # kernels run equally close to their rooflines
        "mfma_peak_ratio_fp32_fp64": PEAK_TFLOPS["mixed_fp32"] / PEAK_TFLOPS["fp64"],
        "graph_sweep": graph_run,
        "speedup_vs_fp64": r["value"] / runs["fp64"]["value"],
        "roofline": {"bound": "mfma", "kernel": r["dominant"], "achieved": r["achieved_tflops"], "peak": peak,
                     "unit": "TFLOP/s", "frac": r["achieved_tflops"] / peak, "traffic": traffic,
                     "per_launch_flops": r["per_realization_flops"] * r["realizations_per_launch"],
                     "avg_launch_ms": r["avg_launch_ms"]},
        "kernel_ms_per_step": r["kernel_ms_per_step"],
        "kernel_ms_per_step_source": "warmup steps, every launch timed (the timed steps time the dominant kernel only)",
        "launch_options": options or "defaults",
        "failed_realizations": r["failed_realizations"],
    }
    if main_prec != "fp64":
        line["accuracy_vs_fp64"] = accuracy(plan, outs["fp64"], outs[main_prec])
    plan.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(params, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
