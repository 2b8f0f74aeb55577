"""GPU parity through the C ABI against the FP64 CPU oracle on identical inputs.

Tolerances (complex128 path): transmission eigenvalues agree to |dT| <= 1e-10 T + 2e-13 sqrt(T T_max)
+ 1e-26 T_max, i.e. 1e-10 relative wherever an FP64 sweep resolves T and |d sigma| <= 1e-13 sigma_max
for the singular values sigma = sqrt(T) of closed channels (two FP64 sweeps of the same potential
differ by that much: tests/test_oracle.py::test_fp64_sweep_resolves_sigma_to_eps_sigma_max); the
Gram eigensolver (option eigen = 0), whose eigenvalues carry ~ n eps T_max, gets an absolute floor of
1e-13 instead; t matrices to 1e-10 relative to max|t|; generated potentials to 4 ulp; histogram
counts bit-exact.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding, distrib

pytestmark = pytest.mark.gpu

RTOL, SIGMA, GRAM_FLOOR = 1e-10, 1e-13, 1e-13


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def T_bound(To, rtol=RTOL, sigma=SIGMA, floor=0.0):
    """Allowed |dT| for the oracle's eigenvalues To: rtol T, plus what a singular-value error
    |d sigma| <= sigma sigma_max does to T = sigma^2, (sqrt(T) + sigma sqrt(T_max))^2 - T, plus `floor`."""
    To = np.abs(np.asarray(To, dtype=float))
    tmax = np.max(To) if To.size else 0.0
    return rtol * To + 2.0 * sigma * np.sqrt(To * tmax) + sigma * sigma * tmax + floor


def assert_T_close(Tg, To, rtol=RTOL, sigma=SIGMA, floor=0.0):
    Tg, To = np.asarray(Tg), np.asarray(To)
    err = np.abs(Tg - To)
    tol = T_bound(To, rtol, sigma, floor)
    bad = err > tol
    assert not bad.any(), f"max rel err {np.max(err / np.maximum(np.abs(To), 1e-300)):.3e}, {bad.sum()} bad of {bad.size}"


def upot_batch(ow, seeds):
    """Oracle potentials stacked as upot(ix, iy, r) (Fortran order)."""
    return np.stack([ow.potential(s).T for s in seeds], axis=2)


CASES = {
    # name: params (reference mode unless ny is given)
    "ref_N10_free": dict(boundtype="dirichlet", nxdiso=100, nxfree=20, wol=4.3, exfac=1.2, dscat=0.2),
    "c1_N20_ny205": dict(boundtype="dirichlet", nxdiso=200, wol=10.25, aratio=10 / 10.25, dscat=2.0, ny=205,
                         dos_model=1),
    "N37_absorb": dict(boundtype="dirichlet", nxdiso=60, wol=18.7, dscat=3.0, dabso=0.5, ny=80),
    "N64": dict(boundtype="dirichlet", nxdiso=40, wol=32.25, aratio=2.0 / 32.25, dscat=1.0, ny=645, dos_model=1),
    "N100_filtered": dict(boundtype="dirichlet", nxdiso=30, wol=50.25, dscat=1.0, napera=0.6, naperb=0.9),
    "N128": dict(boundtype="dirichlet", nxdiso=40, wol=64.25, aratio=2.0 / 64.25, dscat=1.0, ny=1285, dos_model=1),
    # fewer slices than the Shannon limit (k dx > pi): B = mu / sin(mu k dx) is large, slices ill-conditioned
    "N128_coarse": dict(boundtype="dirichlet", nxdiso=30, wol=64.25, aratio=50 / 64.25, dscat=5.0, ny=1285,
                        dos_model=1),
    "periodic_N11": dict(boundtype="periodic", nxdiso=80, nxfree=5, wol=4.3, exfac=1.2, dscat=0.5),
    "periodic_N41_ny60": dict(boundtype="periodic", nxdiso=50, wol=20.3, dscat=2.0, ny=60, napera=0.5),
}


@pytest.mark.parametrize("name", list(CASES))
def test_explicit_potential_parity(name):
    p = CASES[name]
    say(f"explicit potential parity {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=4)
    seeds = [1, 2, 3]
    U = upot_batch(ow, seeds)
    Tg, st, code = plan.tdistrib_potential(U)
    assert code == 0 and not st.any()
    Tg = Tg.reshape(len(seeds), -1)
    for k, s in enumerate(seeds):
        assert_T_close(Tg[k], ow.tdistrib_potential(U[:, :, k].T))
    plan.close()


def zgetrf_rejects_diagonal(M, margin=1e-4):
    """True if partial pivoting on M (max |Re| + |Im| in the column, src/base_utils.f90:235-274) picks an
    off-diagonal row at some step, with the verified kernels' relative margin."""
    A = np.array(M, dtype=complex)
    for j in range(len(A)):
        col = np.abs(A[j:, j].real) + np.abs(A[j:, j].imag)
        if col.size > 1 and col[1:].max() > (1.0 - margin) * col[0]:
            return True
        A[j + 1:, j:] -= np.outer(A[j + 1:, j] / A[j, j], A[j, j:])
    return False


@pytest.mark.parametrize("name", ["N64", "N128"])
def test_pivot_check_fallback(name):
    """A potential spike of 1e5 standard deviations in one cell makes zgetrf pivot off the diagonal: the
    verified kernels must flag those realizations and the pivoted kernel invert them again from the
    unchanged Gss, in the first slice (realization 0) and in slice 20 (realization 1, with B Gss B)."""
    p = CASES[name]
    say(f"pivot-check fallback {name}")
    ow = O.Waveguide(**p)
    U = upot_batch(ow, [4, 5, 6])
    amp = 1e5 * U.std()
    U[0, 10, 0] += amp
    U[19, 10, 1] += amp
    assert zgetrf_rejects_diagonal(ow.slice_matrix(1, U[:, :, 0].T))
    assert zgetrf_rejects_diagonal(ow.slice_matrix(20, U[:, :, 1].T))
    To = np.stack([ow.tdistrib_potential(U[:, :, k].T) for k in range(3)])
    plan = binding.Plan(p, batch_max=3)
    for v in (3, 4):
        plan.set_option("inverse_fp64", v)
        Tg, st, code = plan.tdistrib_potential(U, "fp64")
        assert code == 0 and not st.any(), v
        for k in range(3):
            assert_T_close(Tg.reshape(3, -1)[k], To[k])
    for v in (3, 4):
        plan.set_option("inverse_mixed", v)
        Tg, st, code = plan.tdistrib_potential(U, "mixed_fp32")
        assert code == 0 and not st.any(), v
        err = np.max(np.abs(Tg.reshape(3, -1) - To))
        say(f"{name} mixed (inverse_mixed={v}) with the fallback: max |dT| = {err:.3e}")
        assert err <= MIXED_TOL["mixed_fp32"]
    plan.close()


# Long guides whose transmission eigenvalues span 19 to 20 decades (localized, L / l_s = 25, or strongly
# absorbing), and 7 decades through the filtered apertures (t with fewer columns, or fewer rows, than modes).
CLOSED = {
    "localized_N20": dict(boundtype="dirichlet", nxdiso=1000, wol=10.25, aratio=50 / 10.25, dscat=25.0, ny=205,
                          dos_model=1),
    "localized_tall": dict(boundtype="dirichlet", nxdiso=1000, wol=10.25, aratio=50 / 10.25, dscat=25.0, ny=205,
                           dos_model=1, napera=0.6),
    "localized_wide": dict(boundtype="dirichlet", nxdiso=1000, wol=10.25, aratio=50 / 10.25, dscat=25.0, ny=205,
                           dos_model=1, naperb=0.6),
    "localized_periodic": dict(boundtype="periodic", nxdiso=1000, wol=10.3, aratio=50 / 10.3, dscat=25.0, ny=206),
    "absorbing_N64": dict(boundtype="dirichlet", nxdiso=600, wol=32.125, aratio=30 / 32.125, dscat=3.0, dabso=10.0,
                          ny=643, dos_model=1),
}


@pytest.mark.parametrize("name", list(CLOSED))
def test_closed_channels_relative_accuracy(name):
    """The default complex128 path (singular values of t by bidiagonalisation) and the Jacobi SVD keep
    closed channels to |d sigma| <= 1e-13 sigma_max with no absolute floor; the eigenvalues of t^H t
    (option eigen = 0) are printed for comparison: they carry ~ eps T_max absolute."""
    p = CLOSED[name]
    ow = O.Waveguide(**p)
    U = upot_batch(ow, [3, 4])
    To = [ow.tdistrib_potential(U[:, :, k].T) for k in range(2)]
    assert min(t.max() / t.min() for t in To) > 1e7
    plan = binding.Plan(p, batch_max=2)
    for eigen in (3, 2, 1, 0):
        plan.set_option("eigen", eigen)
        Tg, st, code = plan.tdistrib_potential(U)
        assert code == 0 and not st.any()
        Tg = Tg.reshape(2, -1)
        ratio = max(np.max(np.abs(Tg[k] - To[k]) / T_bound(To[k])) for k in range(2))
        res = [To[k] >= 1e-11 * To[k].max() for k in range(2)]
        rel = max(np.max(np.abs(Tg[k] - To[k])[res[k]] / To[k][res[k]]) for k in range(2))
        say(f"closed channels {name} eigen={eigen}: T down to {min(t.min() for t in To):.1e}, max rel err where "
            f"resolved {rel:.1e}, max err/bound {ratio:.2e}")
        if eigen != 0:
            for k in range(2):
                assert_T_close(Tg[k], To[k])
    plan.close()


@pytest.mark.parametrize("name", ["ref_N10_free", "c1_N20_ny205", "N37_absorb", "periodic_N41_ny60"])
def test_transmission_matrix_parity(name):
    p = CASES[name]
    say(f"t-matrix parity {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=2)
    U = upot_batch(ow, [7, 8])
    tg, st, code = plan.transmission_potential(U)
    assert code == 0
    for k in range(2):
        to = ow.transmission_matrix(U[:, :, k].T)
        assert np.max(np.abs(tg[k] - to)) <= RTOL * np.max(np.abs(to))
    plan.close()


@pytest.mark.parametrize("name", ["ref_N10_free", "c1_N20_ny205", "periodic_N11"])
def test_seeded_generator_and_batch_parity(name):
    """The device generator reproduces the oracle's Philox/Box-Muller field, and the seeded batch
    (rg_tdistrib, spanning several device passes) equals the oracle's OpenMP loop."""
    p = CASES[name]
    say(f"seeded parity {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=5)
    Ug = plan.generate_potential(11, 3)
    for k in range(3):
        Uo = ow.potential(11 + k).T
        np.testing.assert_allclose(Ug[:, :, k], Uo, rtol=4e-16 * 4, atol=4e-16 * 4 * ow.stdu)
    nseed = 12
    Tg, st, code = plan.tdistrib(11, nseed)
    assert code == 0 and not st.any()
    To = ow.tdistrib_batch(11, nseed, nthreads=4)
    assert_T_close(Tg, To)
    plan.close()


@pytest.mark.parametrize("boundtype", ["dirichlet", "periodic"])
def test_zero_disorder_closed_form(boundtype):
    p = dict(boundtype=boundtype, nxdiso=40, nxfree=3, wol=12.3, dscat=0.0, dabso=0.7)
    say(f"zero disorder {boundtype}")
    plan = binding.Plan(p, batch_max=1)
    ow = O.Waveguide(**p)
    T, st, code = plan.transmission_potential(None)
    t = T[0]
    th = ow.mu * ow.kdx
    ref = np.exp(1j * th * (ow.nx - 1)) * 2j * ow.mu / (2j * ow.mu)
    np.testing.assert_allclose(np.diag(t), ref[ow.imin - 1:ow.imax], rtol=1e-11)
    plan.close()


def test_histogram_counts_bit_exact():
    say("histogram counts")
    plan = binding.Plan(CASES["ref_N10_free"], batch_max=1)
    rng = np.random.default_rng(3)
    s = np.concatenate([rng.random(200000), [0.0, 1.0, -1e-9, 1.0 + 1e-12, np.nan, np.nan, 0.5]])
    for bt, nb in (("chebyshev", 256), ("linear", 50), ("chebyshev", 7)):
        e = distrib.bins(bt, nb)
        e[0] = 0.0
        c, miss, drop = plan.histogram(s, e)
        h, m = O.histogram(s, e)
        np.testing.assert_array_equal(c, h.astype(np.int64))
        assert miss == m and drop == 2
        # exact edge values land where the reference's linear scan puts them
        c2, _, _ = plan.histogram(e.copy(), e)
        h2, _ = O.histogram(e, e)
        np.testing.assert_array_equal(c2, h2.astype(np.int64))
    plan.close()


def test_nonfinite_potential_rejected():
    say("non-finite potential")
    p = CASES["ref_N10_free"]
    plan = binding.Plan(p, batch_max=2)
    U = np.zeros((plan.nxdiso, plan.ny, 1), order="F")
    U[3, 2, 0] = np.nan
    T, st, code = plan.tdistrib_potential(U)
    assert code == binding.RG_ENONFINITE
    plan.close()


# ---------------------------------------------------------------------------------------------
# Mixed precision (complex64 state, FP64 pivoting / projection / singular values).
# Errors are stated against the FP64 oracle on identical potentials; bounds are absolute on T.
# ---------------------------------------------------------------------------------------------
MIXED_TOL = {"mixed_fp32": 2e-4, "mixed_bf16x3": 5e-3}


@pytest.mark.parametrize("precision", ["mixed_fp32", "mixed_bf16x3"])
@pytest.mark.parametrize("name", ["ref_N10_free", "c1_N20_ny205", "N64", "N128", "periodic_N41_ny60"])
def test_mixed_precision_against_oracle(name, precision):
    p = CASES[name]
    say(f"mixed {precision} {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=4)
    U = upot_batch(ow, [1, 2, 3])
    Tg, st, code = plan.tdistrib_potential(U, precision)
    assert code == 0 and not st.any()
    Tg = Tg.reshape(3, -1)
    To = np.stack([ow.tdistrib_potential(U[:, :, k].T) for k in range(3)])
    err = np.max(np.abs(Tg - To))
    print(f"[gpu-case] {precision} {name}: max |dT| = {err:.3e}, mean dT = {np.mean(Tg - To):.3e}", flush=True)
    assert err <= MIXED_TOL[precision]
    plan.close()


# ---------------------------------------------------------------------------------------------
# Reflection and transmission matrices (compute_refl_trans_v1, complex128).
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["ref_N10_free", "c1_N20_ny205", "N37_absorb", "periodic_N41_ny60", "N64"])
def test_refl_trans_parity(name):
    p = CASES[name]
    say(f"refl/trans parity {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=3)
    U = upot_batch(ow, [5, 6, 7])
    r, t, st, code = plan.refl_trans(U)
    assert code == 0 and not st.any()
    for k in range(3):
        ro, to = ow.refl_trans(U[:, :, k].T, version=1)
        assert np.max(np.abs(t[k] - to)) <= RTOL * max(1.0, np.max(np.abs(to)))
        assert np.max(np.abs(r[k] - ro)) <= RTOL * max(1.0, np.max(np.abs(ro)))
    plan.close()


def test_refl_trans_unitarity_seeded_n128():
    """Flux conservation r^H r + t^H t = I at zero absorption (src/waveguide.test.f90:54-86) on the
    N=128 workload shape, over more realizations than one device pass holds."""
    p = dict(CASES["N128"], nxdiso=200, aratio=10.0 / 64.25)
    say("refl/trans unitarity N128")
    plan = binding.Plan(p, batch_max=300)
    r, t, st, code = plan.refl_trans(None, seed0=3, nseed=260)
    assert code == 0 and not st.any()
    eye = np.eye(plan.nmu)
    for k in (0, 127, 255, 259):
        err = np.max(np.abs(r[k].conj().T @ r[k] + t[k].conj().T @ t[k] - eye))
        assert err < 1e-9, err
    plan.close()
