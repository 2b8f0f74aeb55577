"""Parity at the full sizes of the benchmark configurations, against the FP64 oracle.

The oracle's answers for the long guides are committed fixtures (tests/golden/fullsize_oracle.json,
written by tests/golden/make_fullsize.py from the same seeds: the device generator reproduces the
oracle's potentials to a few ulp); the shorter ones (the complete c1 run, the c4 adjoints) run the
oracle here.

Bounds.  An FP64 sweep resolves a singular value sigma of t only to ~ eps sigma_max: two FP64
implementations of the same sweep (the oracle against itself with another summation order, or the
reference's zgesvd against a Jacobi SVD of the same t) differ by that much, which is a large
relative error for closed channels (tests/test_oracle.py::test_fp64_sweep_resolves_sigma_to_eps_sigma_max).
The FP64 bar is therefore 1e-10 relative wherever the sweep resolves T (T >~ 1e-11 T_max) and
|d sigma| <= 1e-13 sigma_max below, i.e. |dT| <= 1e-10 T + 2e-13 sqrt(T T_max) + 1e-26 T_max, with no
absolute floor.  The complex64 state resolves sigma to ~ eps32 sigma_max times the growth over the
slices: mixed results are held to |d sigma| <= 1e-4 sigma_max, |dT| <= 2e-4 sqrt(T T_max) + 1e-8 T_max,
and the mean of T to 2e-6.
"""
import concurrent.futures as cf
import json
import os

import numpy as np
import pytest

from fullsize_cases import UNIFORM, guide, uneven_sample
from oracle import oracle as O
from recurgreen_amd import binding, distrib

from test_gpu_parity import T_bound

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_oracle.json")))


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def fp64_bound(To):
    return T_bound(To)  # 1e-10 T + 2e-13 sqrt(T T_max) + 1e-26 T_max


def mixed_bound(To):
    return T_bound(To, rtol=0.0, sigma=1e-4)  # 2e-4 sqrt(T T_max) + 1e-8 T_max


def fixture(name):
    ent = FIX[name]
    return ent[0]["params"], [e["seed"] for e in ent], [np.array(e["T"]) for e in ent]


def check(Tg, To, bound, label, strict=True):
    """Per-realization check; prints the largest relative error where the sweep resolves T, the
    largest singular-value error relative to sigma_max and the largest ratio of error to bound."""
    err = np.abs(np.asarray(Tg) - To)
    tol = bound(To)
    res = To >= 1e-11 * np.max(To)
    rel = np.max(err[res] / To[res]) if res.any() else 0.0
    sig = np.max(np.abs(np.sqrt(np.abs(np.asarray(Tg))) - np.sqrt(To))) / np.sqrt(np.max(To))
    say(f"{label}: T in [{To.min():.2e}, {To.max():.2e}], max rel err (T >= 1e-11 T_max) {rel:.2e}, "
        f"max |d sigma|/sigma_max {sig:.2e}, max err/bound {np.max(err / tol):.2e}")
    ok = bool(np.all(err <= tol))
    assert ok or not strict, f"{label}: {np.sum(err > tol)} of {err.size} outside the bound"
    return rel, sig, ok


@pytest.mark.parametrize("name", list(UNIFORM))
def test_fullsize_fp64_against_oracle(name):
    p, seeds, Ts = fixture(name)
    assert seeds == list(range(seeds[0], seeds[0] + len(seeds)))
    say(f"full-size fp64 {name}: N={int(2 * p['wol'])} nx={p['nxdiso']} ny={p['ny']}")
    plan = binding.Plan(p, batch_max=len(seeds))
    T, st, code = plan.tdistrib(seeds[0], len(seeds), "fp64")
    plan.close()
    assert code == 0 and not st.any()
    T = T.reshape(len(seeds), -1)
    for k, s in enumerate(seeds):
        check(T[k], Ts[k], fp64_bound, f"{name} seed {s} fp64")


@pytest.mark.parametrize("name", ["n128", "c3_nx200", "c3"])
def test_fullsize_mixed_against_oracle(name):
    p, seeds, Ts = fixture(name)
    say(f"full-size mixed_fp32 {name}: N={int(2 * p['wol'])} nx={p['nxdiso']} ny={p['ny']}")
    plan = binding.Plan(p, batch_max=len(seeds))
    T, st, code = plan.tdistrib(seeds[0], len(seeds), "mixed_fp32")
    plan.close()
    assert code == 0 and not st.any()
    T = T.reshape(len(seeds), -1)
    To = np.stack(Ts)
    for k, s in enumerate(seeds):
        check(T[k], Ts[k], mixed_bound, f"{name} seed {s} mixed_fp32")
    shift = float(T.mean() - To.mean())
    say(f"{name} mixed_fp32: max |dT| {np.max(np.abs(T - To)):.2e}, mean T shift {shift:+.2e} over {To.size} values")
    assert abs(shift) <= 2e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed_fp32"])
def test_fullsize_uneven_against_oracle(precision):
    """The first two realizations of each width of the 10000-realization uneven workload (N in {16..256},
    W = N/2 + 1/8, ny = 20 W, L/l_s in [0.5, 10], eps in [0, 0.05 k^2]) in one rg_tdistrib_uneven
    call, each against the oracle; errors reported per N."""
    ps, seeds = uneven_sample()
    fx = {e["seed"]: np.array(e["T"]) for e in FIX["c5"]}
    say(f"full-size uneven batch of {len(ps)} ({precision})")
    vals, st, code = binding.tdistrib_uneven(ps, np.array(seeds, np.int64), precision, batch_max=8)
    assert code == 0 and not st.any()
    per_n, bad = {}, []
    bound = fp64_bound if precision == "fp64" else mixed_bound
    for p, s, T in zip(ps, seeds, vals):
        n = int(2 * p["wol"])
        label = f"uneven N={n} nx={p['nxdiso']} dabso={p['dabso']:.2f} seed {s} {precision}"
        rel, sig, ok = check(T, fx[s], bound, label, strict=False)
        if not ok:
            bad.append(label)
        a, b = per_n.get(n, (0.0, 0.0))
        per_n[n] = (max(a, rel), max(b, sig))
    say(f"uneven {precision} per N (max rel err where resolved, max |d sigma|/sigma_max): "
        + ", ".join(f"N={n}: {a:.1e}/{b:.1e}" for n, (a, b) in sorted(per_n.items())))
    assert not bad, bad


def test_c1_complete_run_against_oracle():
    """c1 in full: 1000 realizations (seeds 1..1000) of W = 10.25 (N = 20), L = 10 on 200 slices, ny = 205,
    l_s = 5, every eigenvalue against the oracle's OpenMP loop, and the 50-bin rho(T) histograms (Chebyshev
    bins as the tdistrib task, and linear) counted on the device equal to the oracle's counts."""
    p = guide(10.25, 200, 5.0)
    say("c1 complete: 1000 realizations N=20 nx=200")
    plan = binding.Plan(p, batch_max=1000)
    T, st, code = plan.tdistrib(1, 1000, "fp64")
    assert code == 0 and not st.any()
    To = O.Waveguide(**p).tdistrib_batch(1, 1000, nthreads=min(16, os.cpu_count() or 1))
    T, To = T.reshape(1000, -1), To.reshape(1000, -1)
    for k in range(1000):
        assert np.all(np.abs(T[k] - To[k]) <= fp64_bound(To[k])), k
    for bt in ("chebyshev", "linear"):
        e = distrib.bins(bt, 50)
        cg, mg, dg = plan.histogram(T.ravel(), e)
        co, mo = O.histogram(To.ravel(), e)
        assert np.array_equal(cg, co.astype(np.int64)) and mg == mo and dg == 0, bt
        rho_g = distrib.rho_from_counts(cg, e)
        rho_o, _ = O.prepare_distrib(To.ravel(), 50, bt)
        np.testing.assert_allclose(rho_g, rho_o, rtol=1e-13)
    plan.close()


C4 = guide(32.25, 400, 10.0)


def test_c4_tmax_gradient_finite_differences():
    """c4 (W = 32.25, N = 64, 400 slices, ny = 645, l_s = 10): the adjoint gradient of T_max against
    central differences of the device's FP64 T_max on 64 random cells (step 1e-6) and on 16 cells with
    the fourth-order stencil of step 1e-3, each to 1e-6 relative plus the differences' own round-off:
    each FP64 T_max carries eps_T ~ 3e-14 T_max of rounding after 400 slices (~ sqrt(400) eps per
    slice; printed below as the difference between two inversion kernels), so a central difference
    carries eps_T / h, which dominates at step 1e-6 where a cell's derivative is far below the largest."""
    say("c4 T_max gradient vs finite differences")
    plan = binding.Plan(C4, batch_max=130)
    U = plan.generate_potential(17, 1)[:, :, 0]
    val, grad, st, code = plan.adjoint(U, "tmax")
    assert code == 0 and not st.any()
    T0 = plan.tdistrib_potential(U)[0]
    assert val[0] == pytest.approx(T0[0], rel=1e-12)
    rng = np.random.default_rng(11)

    def tmax(perturb):
        batch = np.repeat(U[:, :, None], len(perturb), axis=2)
        for k, (ix, iy, dh) in enumerate(perturb):
            batch[ix, iy, k] += dh
        T, st2, code2 = plan.tdistrib_potential(batch)
        assert code2 == 0 and not st2.any()
        return T.reshape(len(perturb), -1)[:, 0]

    eps_t = 3e-14 * val[0]
    plan.set_option("inverse_fp64", 3)  # the same T_max through the general inversion: its rounding differs
    noise = abs(plan.tdistrib_potential(U)[0][0] - T0[0])
    plan.set_option("inverse_fp64", 4)
    assert noise <= eps_t
    cells = [(int(rng.integers(C4["nxdiso"])), int(rng.integers(C4["ny"]))) for _ in range(64)]
    h = 1e-6
    L = tmax([(ix, iy, s * h) for ix, iy in cells for s in (1, -1)])
    floor = eps_t / h
    worst = 0.0
    for c, (ix, iy) in enumerate(cells):
        fd = (L[2 * c] - L[2 * c + 1]) / (2 * h)
        err = abs(grad[ix, iy, 0] - fd)
        worst = max(worst, err / (abs(fd) + floor))
        assert err <= 1e-6 * abs(fd) + floor, (ix, iy, grad[ix, iy, 0], fd)
    h = 1e-3
    L4 = tmax([(ix, iy, s * h) for ix, iy in cells[:16] for s in (2, 1, -1, -2)])
    floor4 = 1.5 * eps_t / h  # round-off of the stencil: weights (1, 8, 8, 1) / 12h on eps_t
    worst4 = 0.0
    for c, (ix, iy) in enumerate(cells[:16]):
        f2, f1, m1, m2 = L4[4 * c:4 * c + 4]
        fd4 = (-f2 + 8 * f1 - 8 * m1 + m2) / (12 * h)
        err = abs(grad[ix, iy, 0] - fd4)
        worst4 = max(worst4, err / (abs(fd4) + floor4))
        assert err <= 1e-6 * abs(fd4) + floor4, (ix, iy, grad[ix, iy, 0], fd4)
    say(f"c4 T_max: step 1e-6 max err/(|fd|+floor) {worst:.2e} (floor {floor:.1e}); 4th order step 1e-3 max err/(|fd|+floor) "
        f"{worst4:.2e} (floor {floor4:.1e}); T1 {T0[0]:.6f}, gap {T0[0] - T0[1]:.2e}; T_max rounding between "
        f"inversion kernels {noise / val[0]:.1e} relative")
    plan.close()


@pytest.mark.parametrize("loss", ["g", "tmax"])
def test_c4_adjoint_against_cpu_adjoint(loss):
    """c4: the device's complex128 adjoint of both losses against the FP64 CPU adjoint on the same two
    potentials, loss to 1e-10 and gradient to 1e-9 of max |grad| (g) or 1e-8 (T_max, through the top
    eigenvector)."""
    say(f"c4 adjoint {loss} vs oracle")
    plan = binding.Plan(C4, batch_max=2)
    ow = O.Waveguide(**C4)
    U = np.stack([ow.potential(s).T for s in (3, 4)], axis=2)
    val, grad, st, code = plan.adjoint(U, loss)
    assert code == 0 and not st.any()
    with cf.ThreadPoolExecutor(2) as ex:  # ctypes releases the GIL during oracle calls
        ref = list(ex.map(lambda k: ow.adjoint_grad(U[:, :, k].T, loss), range(2)))
    for k, (vo, go) in enumerate(ref):
        assert val[k] == pytest.approx(vo, rel=1e-10)
        err = np.max(np.abs(grad[:, :, k].T - go)) / np.max(np.abs(go))
        say(f"c4 {loss} realization {k}: loss {val[k]:.12g}, gradient err {err:.2e} of max |grad|")
        assert err <= (1e-9 if loss == "g" else 1e-8)
    plan.close()
