"""FP64 refinement of G(L;0) in the complex64 sweep (plan option refine = k), against the FP64 oracle on the
same potentials.

The refined sweep keeps every slice's complex64 inverse g_x and corrects X = H^-1 E_1 by the FP64 residual
E_1 - H X solved with them (recurgreen_amd/csrc/rg_refine.hip); G(L;0) = X_nx then goes through the FP64
Fisher-Lee relation and eigensolver.  Bar: the complex128 path's (test_gpu_parity.T_bound: 1e-10 relative
wherever an FP64 sweep resolves T, |d sigma| <= 1e-13 sigma_max below), where the plain complex64 sweep
misses it by orders of magnitude (its ratio of error to bound is printed beside).  Cases: absorbing, N = 64
and N = 128 Dirichlet, a periodic guide with filtered apertures, and localized guides of 1000 slices whose
eigenvalues span 19 to 20 decades (three passes there, where each pass gains about 2e5; two elsewhere);
at full size, the N = 128 workload (1000 slices, ny = 1285, l_s = 10)
and N = 256 over 200 slices against the committed oracle fixture.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding

from test_gpu_parity import CASES, CLOSED, RTOL, T_bound, assert_T_close, say, upot_batch

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_oracle.json")))


def worst_ratio(Tg, To):
    return float(np.max(np.abs(np.asarray(Tg) - To) / T_bound(To)))


@pytest.mark.parametrize("name", ["N37_absorb", "N64", "N128", "periodic_N41_ny60", "localized_N20",
                                  "localized_periodic"])
def test_refined_sweep_matches_fp64_oracle(name):
    p = CASES[name] if name in CASES else CLOSED[name]
    say(f"refined mixed sweep {name}")
    ow = O.Waveguide(**p)
    seeds = [11, 12, 13]
    U = upot_batch(ow, seeds)
    To = [ow.tdistrib_potential(U[:, :, k].T) for k in range(len(seeds))]
    plan = binding.Plan(p, batch_max=4)
    plain, _, _ = plan.tdistrib_potential(U, "mixed_fp32")
    passes = 3 if name.startswith("localized") else 2
    plan.set_option("refine", passes)
    assert plan.get_option("refine") == passes
    Tg, st, code = plan.tdistrib_potential(U, "mixed_fp32")
    assert code == 0 and not st.any()
    Tg, plain = Tg.reshape(len(seeds), -1), plain.reshape(len(seeds), -1)
    for k in range(len(seeds)):
        say(f"{name} seed {seeds[k]}, {passes} passes: error / FP64 bound, complex64 sweep {worst_ratio(plain[k], To[k]):.2e}, "
            f"refined {worst_ratio(Tg[k], To[k]):.2e}")
        assert_T_close(Tg[k], To[k])
    # complex128 sweeps do not take the option
    f1, _, _ = plan.tdistrib_potential(U, "fp64")
    plan.set_option("refine", 0)
    f0, _, _ = plan.tdistrib_potential(U, "fp64")
    assert np.array_equal(f1, f0)
    plan.close()


@pytest.mark.parametrize("passes", [2, 3])
def test_refined_transmission_matrix(passes):
    p = CASES["N64"]
    say(f"refined t matrix N64, {passes} passes")
    ow = O.Waveguide(**p)
    U = upot_batch(ow, [7, 8])
    plan = binding.Plan(p, batch_max=2)
    plan.set_option("refine", passes)
    tg, st, code = plan.transmission_potential(U, precision="mixed_fp32")
    assert code == 0 and not st.any()
    for k in range(2):
        to = ow.transmission_matrix(U[:, :, k].T)
        err = np.max(np.abs(tg[k] - to)) / np.max(np.abs(to))
        say(f"N64 realization {k}: max |dt| / max |t| = {err:.2e}")
        assert err <= RTOL
    plan.close()


def test_refined_passes_split_by_refine_cap():
    """Five realizations through passes of 4 + 1 (batch_max) and of 2 + 2 + 1 (refine_cap = 2): each within
    the FP64 bar of the oracle either way."""
    p = CASES["N64"]
    say("refined passes split by refine_cap, N64")
    ow = O.Waveguide(**p)
    seeds = [21, 22, 23, 24, 25]
    U = upot_batch(ow, seeds)
    To = [ow.tdistrib_potential(U[:, :, k].T) for k in range(len(seeds))]
    plan = binding.Plan(p, batch_max=4)
    plan.set_option("refine", 2)
    out = {}
    for cap in (0, 2):
        plan.set_option("refine_cap", cap)
        assert plan.get_option("refine_cap") == cap
        T, st, code = plan.tdistrib_potential(U, "mixed_fp32")
        assert code == 0 and not st.any()
        out[cap] = T.reshape(len(seeds), -1)
        for k in range(len(seeds)):
            assert_T_close(out[cap][k], To[k])
    with pytest.raises(binding.RecurGreenError):
        plan.set_option("refine_cap", -1)
    plan.close()
    say(f"N64 passes 4 + 1 against 2 + 2 + 1: max |dT| {np.max(np.abs(out[0] - out[2])):.2e}")


@pytest.mark.parametrize("name", ["n128", "c3_nx200"])
def test_refined_fullsize_against_oracle(name):
    ent = FIX[name]
    p, seeds, Ts = ent[0]["params"], [e["seed"] for e in ent], [np.array(e["T"]) for e in ent]
    assert seeds == list(range(seeds[0], seeds[0] + len(seeds)))
    say(f"refined full size {name}: N={int(2 * p['wol'])} nx={p['nxdiso']} ny={p['ny']}")
    plan = binding.Plan(p, batch_max=len(seeds))
    plan.set_option("refine", 2)
    T, st, code = plan.tdistrib(seeds[0], len(seeds), "mixed_fp32")
    plan.close()
    assert code == 0 and not st.any()
    T = T.reshape(len(seeds), -1)
    for k, s in enumerate(seeds):
        res = Ts[k] >= 1e-11 * Ts[k].max()
        rel = np.max(np.abs(T[k][res] - Ts[k][res]) / Ts[k][res])
        say(f"{name} seed {s}: max rel err (T >= 1e-11 T_max) {rel:.2e}, error / bound {worst_ratio(T[k], Ts[k]):.2e}")
        assert_T_close(T[k], Ts[k])
