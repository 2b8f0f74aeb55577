"""Launch options and the autotuner (no counterpart in the reference).

Every option value runs the same sweep with a different kernel or launch shape, so each must give
the oracle's answer: complex128 sweeps within the bound of tests/test_gpu_parity.py (the Gram
eigensolver with its absolute floor of 1e-13), complex64 sweeps within
the absolute bound of the mixed-precision tests; the two-lane batch split must be bit-identical
to one stream (each realization sees the same kernels and data).
"""
import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import autotune, binding

from test_gpu_parity import CASES, GRAM_FLOOR, MIXED_TOL, assert_T_close, say, upot_batch

pytestmark = pytest.mark.gpu


def test_option_round_trip_and_rejection():
    say("options round trip")
    plan = binding.Plan(CASES["c1_N20_ny205"], batch_max=2)
    for key, values in (("inverse_fp64", (0, 1, 2, 3, 4)), ("inverse_mixed", (1, 2, 3, 4)), ("dyson", (0, 1, 2, 3)),
                        ("project_parts", (1, 3, 8)), ("project_wide_min", (4096, 1)), ("refine", (2, 8, 0)),
                        ("lanes", (2, 1)),
                        ("eigen", (1, 0, 2, 3))):
        for v in values:
            plan.set_option(key, v)
            assert plan.get_option(key) == v
    for key, v in (("inverse_fp64", 5), ("inverse_mixed", 0), ("dyson", 4), ("project_parts", 9),
                   ("project_wide_min", 0), ("refine", 9), ("refine", -1), ("lanes", 3), ("eigen", 4), ("nonsense", 0)):
        with pytest.raises(binding.RecurGreenError):
            plan.set_option(key, v)
    plan.close()


@pytest.mark.parametrize("name", ["N37_absorb", "N100_filtered", "N128", "periodic_N41_ny60"])
def test_every_option_value_matches_oracle(name):
    p = CASES[name]
    say(f"option values against the oracle {name}")
    ow = O.Waveguide(**p)
    U = upot_batch(ow, [4, 5])
    To = np.stack([ow.tdistrib_potential(U[:, :, k].T) for k in range(2)])
    plan = binding.Plan(p, batch_max=2)
    for key, values in (("inverse_fp64", (0, 1, 2, 3, 4)), ("dyson_fp64", (1, 2, 0)), ("project_parts", (1, 2, 4, 8)),
                        ("project_wide_min", (1 << 30, 1)), ("eigen", (1, 0, 2, 3))):
        for v in values:
            plan.set_option(key, v)
            Tg, st, code = plan.tdistrib_potential(U, "fp64")
            assert code == 0 and not st.any(), (key, v)
            gram = key == "eigen" and v == 0  # the Gram eigensolver: ~ n eps T_max absolute
            for k in range(2):
                assert_T_close(Tg.reshape(2, -1)[k], To[k], floor=GRAM_FLOOR if gram else 0.0)
    for key, values in (("inverse_mixed", (1, 2, 3, 4)), ("dyson", (0, 1, 2, 3)), ("eigen", (2, 1, 0, 3))):
        for v in values:
            plan.set_option(key, v)
            Tg, st, code = plan.tdistrib_potential(U, "mixed_fp32")
            assert code == 0 and not st.any(), (key, v)
            err = np.max(np.abs(Tg.reshape(2, -1) - To))
            print(f"[gpu-case] {name} mixed {key}={v}: max |dT| = {err:.3e}", flush=True)
            assert err <= MIXED_TOL["mixed_fp32"], (key, v)
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed_fp32"])
def test_two_lanes_bit_identical(precision):
    import torch

    p = dict(CASES["N64"], nxdiso=12)
    say(f"two lanes vs one stream {precision}")
    B = 300  # two lanes need at least 2 x 128 realizations
    plan = binding.Plan(p, batch_max=B)
    outs = []
    for lanes in (1, 2):
        plan.set_option("lanes", lanes)
        out = torch.empty(B * plan.neigv, dtype=torch.float64, device="cuda")
        st = torch.zeros(B, dtype=torch.int32, device="cuda")
        plan.tdistrib_device(11, B, precision, out, st)
        plan.synchronize()
        assert int(st.abs().sum()) == 0
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    plan.close()


def test_autotune_cache_round_trip(tmp_path):
    p = CASES["c1_N20_ny205"]
    say("autotune tiny shape")
    path = str(tmp_path / "tune.json")
    entry = autotune.tune(p, 16, "mixed_fp32", probe_slices=8, reps=1, path=path)
    for key, values in autotune.CANDIDATES["mixed_fp32"]:
        if key in entry:
            assert entry[key] in values
    plan = binding.Plan(p, batch_max=16)
    got = autotune.apply(plan, p, 16, "mixed_fp32", path=path)
    assert got is not None
    assert plan.get_option("inverse_mixed") == entry["inverse_mixed"]
    assert plan.get_option("project_parts") == entry["project_parts"]
    assert autotune.apply(plan, p, 32, "mixed_fp32", path=path) is None  # other shape: not cached
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed_fp32"])
def test_graph_replays_match_direct_launches(precision):
    """Captured sweeps replayed several times, with direct sweeps and synchronisations between the
    replays: every replay equals the direct sweep of the same seeds bit for bit."""
    import torch

    p = dict(CASES["N64"], nxdiso=12)
    say(f"graph replays {precision}")
    B = 64
    plan = binding.Plan(p, batch_max=B)

    out = torch.empty(B * plan.neigv, dtype=torch.float64, device="cuda")  # one output: one graph
    st = torch.zeros(B, dtype=torch.int32, device="cuda")

    def run(seed0, graph):
        plan.set_option("graph", graph)
        out.fill_(-1.0)
        plan.tdistrib_device(seed0, B, precision, out, st)
        plan.synchronize()
        assert int(st.abs().sum()) == 0
        return out.cpu().numpy()

    direct = {s: run(s, 0) for s in (1, 101, 201)}
    for s in (1, 101, 1, 201, 101):
        assert np.array_equal(run(s, 1), direct[s]), s
        assert np.array_equal(run(s + 100 if s < 201 else 1, 0), direct[s + 100 if s < 201 else 1])
    # explicit potentials through the staging buffer
    ow = O.Waveguide(**p)
    U = upot_batch(ow, [3, 4])
    plan.set_option("graph", 0)
    T0, st0, c0 = plan.tdistrib_potential(U, precision)
    plan.set_option("graph", 1)
    for _ in range(3):
        T1, st1, c1 = plan.tdistrib_potential(U, precision)
        assert c1 == 0 and np.array_equal(T1, T0)
    plan.close()


@pytest.mark.parametrize("dyson", [2, 3])
@pytest.mark.parametrize("wol", [40.25, 48.25, 56.25, 64.25])
def test_persistent_dyson_matches_oracle(wol, dyson):
    """Options dyson = 2 (the complex64 Dyson product on the BF16 matrix cores with three-way split
    FP32 operands) and 3 (three real FP32 products per complex product), both built for NP = 96 and 128
    (other orders up to 128 take dyson = 1), at NP = 80, 96, 112, 128, within the complex64 bound.
    (NP 160 and 224, the three-product kernel in blocks of 80 and 112: tests/test_gpu_large.py.)"""
    p = dict(boundtype="dirichlet", nxdiso=40, wol=wol, aratio=2.0 / wol, dscat=1.0, ny=int(20 * wol), dos_model=1)
    say(f"Dyson option {dyson} N={int(2 * wol)}")
    ow = O.Waveguide(**p)
    U = upot_batch(ow, [3, 8, 9])
    To = np.stack([ow.tdistrib_potential(U[:, :, k].T) for k in range(3)])
    plan = binding.Plan(p, batch_max=3)
    plan.set_option("dyson", dyson)
    Tg, st, code = plan.tdistrib_potential(U, "mixed_fp32")
    assert code == 0 and not st.any()
    err = np.max(np.abs(Tg.reshape(3, -1) - To))
    print(f"[gpu-case] dyson={dyson} N={int(2 * wol)}: max |dT| = {err:.3e}", flush=True)
    assert err <= MIXED_TOL["mixed_fp32"]
    plan.set_option("dyson", 1)
    T1, st1, code1 = plan.tdistrib_potential(U, "mixed_fp32")
    assert code1 == 0 and np.max(np.abs(T1 - Tg)) <= MIXED_TOL["mixed_fp32"]
    plan.close()


@pytest.mark.parametrize("dyson", [2, 3])
def test_persistent_dyson_loop(dyson):
    """Options dyson = 2 and 3 launch one workgroup per CU, each looping over realizations blockIdx.x +
    k gridDim.x and prefetching the next realization's first chunk under the current one's last.
    With 2 ncu + 3 realizations every workgroup runs two or three of them: all must equal the FP32
    kernel's (dyson = 1) within the complex64 bound, and realizations from the first, second and
    third pass of the loop must match the oracle."""
    import torch

    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    n = 2 * ncu + 3
    p = dict(CASES["N128"], nxdiso=6, aratio=0.3 / 64.25)
    say(f"Dyson option {dyson} persistent loop, {n} realizations over {ncu} workgroups")
    plan = binding.Plan(p, batch_max=n)
    plan.set_option("dyson", dyson)
    Tb, st, code = plan.tdistrib(1, n, "mixed_fp32")
    assert code == 0 and not st.any()
    plan.set_option("dyson", 1)
    T1, st1, code1 = plan.tdistrib(1, n, "mixed_fp32")
    assert code1 == 0 and not st1.any()
    plan.close()
    neigv = Tb.size // n
    Tb, T1 = Tb.reshape(n, neigv), T1.reshape(n, neigv)
    err = np.max(np.abs(Tb - T1), axis=1)
    print(f"[gpu-case] dyson={dyson} loop: max |dT| against dyson=1 {err.max():.3e} (worst realization {err.argmax()})",
          flush=True)
    assert err.max() <= MIXED_TOL["mixed_fp32"]
    ow = O.Waveguide(**p)
    for r in (0, ncu + 1, 2 * ncu + 2):  # first, second and third realization of workgroups 0, 1, 2
        To = ow.tdistrib_batch(1 + r, 1)
        assert np.max(np.abs(Tb[r] - To)) <= MIXED_TOL["mixed_fp32"], r


WIDE_CASES = {
    # one group of column tiles, seeded disorder generated in the kernel (four partial sums at 4096)
    "N128": (dict(CASES["N128"], nxdiso=6), 4096),
    # two groups of column tiles (513 table columns), disorder from k_noise_slice
    "N256": (dict(boundtype="dirichlet", nxdiso=4, wol=128.25, aratio=0.2 / 128.25, dscat=0.3), 1024),
}


@pytest.mark.parametrize("nb", [64, 1024])
@pytest.mark.parametrize("precision", ["fp64", "mixed_fp32"])
def test_generated_disorder_equals_explicit_bit_for_bit(precision, nb):
    """Seeded sweeps up to 128 modes generate each slice's disorder inside the projection kernel
    (k_project_gen); explicit potentials load it (k_project_wide).  Fed the disorder that
    rg_generate_potential writes, the two must give the same transmission eigenvalues bit for bit: the
    generator's values equal the potential's (one Box-Muller cosine on both paths), and at these batch
    sizes both kernels add the same cell ranges (8 partial sums) in the same order."""
    p = dict(CASES["N128"], nxdiso=5)
    say(f"generated against explicit disorder, N128 x {nb}, {precision}")
    plan = binding.Plan(p, batch_max=nb)
    assert plan.get_option("project_parts") == 8
    T1, st1, c1 = plan.tdistrib(21, nb, precision)
    U = plan.generate_potential(21, nb)
    T2, st2, c2 = plan.tdistrib_potential(U, precision)
    assert c1 == 0 and c2 == 0 and not st1.any() and not st2.any()
    assert np.array_equal(T1, T2)
    plan.close()


@pytest.mark.parametrize("case", list(WIDE_CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed_fp32"])
def test_wide_projection_matches_small_batches(precision, case):
    """The wide projection kernels (seeded: 64 realizations per workgroup with the disorder generated in
    the kernel up to 128 modes, else 128 per workgroup in one or two groups of column tiles; explicit
    potentials loaded) on a large batch, against one-realization sweeps on the 32-realization kernel fed
    by k_noise_slice (project_wide_min raised past the batch).  The same realizations through both must
    agree: complex128 within the FP64 bar, complex64 within the mixed bound (the kernels add the cells in
    different orders and cell ranges)."""
    p, nb = WIDE_CASES[case]
    say(f"wide projection against small batches, {case} x {nb}, {precision}")
    big = binding.Plan(p, batch_max=nb)
    small = binding.Plan(p, batch_max=4)
    small.set_option("project_wide_min", 1 << 30)
    picks = [0, 1, nb // 2 - 1, nb - 1]
    Tb, st, code = big.tdistrib(11, nb, precision)
    assert code == 0 and not st.any()
    Tb = Tb.reshape(nb, -1)
    U = big.generate_potential(11, nb)  # (nxdiso, ny, nb)
    Tbu, st, code = big.tdistrib_potential(U, precision)
    assert code == 0 and not st.any()
    Tbu = Tbu.reshape(nb, -1)
    for k in picks:
        Ts, st, code = small.tdistrib(11 + k, 1, precision)
        assert code == 0 and not st.any()
        Tsu, st, code = small.tdistrib_potential(U[:, :, k:k + 1], precision)
        assert code == 0 and not st.any()
        for a, ref in ((Tb[k], Ts), (Tbu[k], Tsu), (Tb[k], Tsu)):
            if precision == "fp64":
                assert_T_close(a, ref)
            else:
                assert np.max(np.abs(a - ref)) <= MIXED_TOL["mixed_fp32"]
    big.close()
    small.close()
