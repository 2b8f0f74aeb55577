"""Deposition matrix (compute_deposition, src/waveguide.f90:710-745) through the C ABI against the
oracle's block-tridiagonal solve on identical explicit potentials.

Tolerances (complex128): 1e-10 relative to max|dep|.  At fpos = nx the deposition matrix is t and
must equal the reflection/transmission pass's t bit for bit (same kernels, same order)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding
from recurgreen_amd.waveguide import Waveguide

pytestmark = pytest.mark.gpu

CASES = {
    "dirichlet_N10_free": dict(boundtype="dirichlet", nxdiso=60, nxfree=5, wol=4.3, exfac=1.2, dscat=0.5),
    "periodic_N11": dict(boundtype="periodic", nxdiso=40, nxfree=3, wol=4.3, dscat=0.7),
    "N37_absorb": dict(boundtype="dirichlet", nxdiso=50, wol=18.7, dscat=3.0, dabso=0.5, ny=80),
    "N64": dict(boundtype="dirichlet", nxdiso=30, wol=32.25, aratio=2.0 / 32.25, dscat=1.0, ny=645, dos_model=1),
    "periodic_N161": dict(boundtype="periodic", nxdiso=12, wol=80.3, aratio=0.1, dscat=1.0),
}


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def upot_batch(ow, seeds):
    return np.stack([ow.potential(s).T for s in seeds], axis=2)


@pytest.mark.parametrize("name", list(CASES))
def test_deposition_matches_oracle(name):
    p = CASES[name]
    say(f"deposition parity {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=3)
    U = upot_batch(ow, [2, 3, 4])
    nx = plan.nx
    mid = p.get("nxfree", 0) + p["nxdiso"] // 2 + 1
    pairs = {(f, 1) for f in (1, 2, mid, nx - 1, nx)}
    # interior and right-hand sources: ipos below fpos, above it, equal to it, and at both ends
    pairs |= {(mid, 3), (3, mid), (mid, mid), (nx, nx), (1, nx), (nx, nx - 1), (1, 2), (2, 2), (nx - 1, nx)}
    for fpos, ipos in sorted(pairs):
        d, st, code = plan.deposition(fpos, ipos, U)
        assert code == 0 and not st.any()
        for k in range(3):
            ref = ow.deposition(U[:, :, k].T, fpos, ipos)
            assert np.max(np.abs(d[k] - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref))), (fpos, ipos, k)
    plan.close()


@pytest.mark.parametrize("boundtype", ["dirichlet", "periodic"])
def test_deposition_reciprocity_n128(boundtype):
    """H is complex symmetric, so G(f; i) = G(i; f)^T: dep(f, i) = dep(i, f)^T for real (Dirichlet)
    modes, with the channel order reversed (k -> -k) for plane waves; on the N = 128 shape, where
    the oracle is slow, for a source left of, right of and at the observation slice."""
    if boundtype == "dirichlet":
        p = dict(boundtype="dirichlet", nxdiso=200, wol=64.25, aratio=10.0 / 64.25, dscat=1.0, ny=1285, dos_model=1)
    else:
        p = dict(boundtype="periodic", nxdiso=200, wol=64.3, aratio=0.15, dscat=1.0, dabso=0.1)
    say(f"deposition reciprocity N128 {boundtype}")
    plan = binding.Plan(p, batch_max=4)
    n = plan.nmu
    perm = np.arange(n) if boundtype == "dirichlet" else np.arange(n)[::-1]
    for f, i in [(150, 40), (7, 180), (plan.nx, 1)]:
        a, st, code = plan.deposition(f, i, None, seed0=11, nseed=4)
        b, st2, code2 = plan.deposition(i, f, None, seed0=11, nseed=4)
        assert code == 0 and code2 == 0 and not st.any() and not st2.any()
        bt = np.swapaxes(b, 1, 2)[:, perm][:, :, perm]
        scale = np.max(np.abs(a))
        assert np.max(np.abs(a - bt)) <= 1e-10 * scale, (f, i)
    plan.close()


def test_deposition_at_the_exit_is_t_bit_for_bit():
    p = CASES["N37_absorb"]
    say("deposition fpos = nx equals t")
    plan = binding.Plan(p, batch_max=8)
    d, st, code = plan.deposition(plan.nx, 1, None, seed0=5, nseed=8)
    r, t, st2, code2 = plan.refl_trans(None, seed0=5, nseed=8)
    assert code == 0 and code2 == 0
    np.testing.assert_array_equal(d, t)
    plan.close()


def test_seeded_deposition_equals_explicit():
    """Seeded passes (more realizations than one pass holds) equal explicit potentials from the
    same generator, and the single-realization host call equals its column of the batch."""
    p = dict(boundtype="dirichlet", nxdiso=80, wol=32.25, aratio=4.0 / 32.25, dscat=2.0, ny=645, dos_model=1)
    say("deposition seeded N64")
    wg = Waveguide(**p, batch_max=300)
    plan = wg.plan
    mid = p["nxdiso"] // 2 + 1
    d, st, code = plan.deposition(mid, 1, None, seed0=1, nseed=260)
    assert code == 0 and not st.any()
    U = plan.generate_potential(255, 3)
    de, st, code = plan.deposition(mid, 1, U)
    np.testing.assert_array_equal(de, d[254:257])
    np.testing.assert_array_equal(wg.compute_deposition(7, mid), d[6])
    wg.delete()


def test_deposition_argument_checks():
    p = CASES["dirichlet_N10_free"]
    say("deposition argument checks")
    plan = binding.Plan(p, batch_max=2)
    with pytest.raises(binding.RecurGreenError) as e:
        plan.deposition(plan.nx + 1, 1, None, nseed=1)
    assert e.value.code == binding.RG_EINVAL
    for ipos in (0, plan.nx + 1):
        with pytest.raises(binding.RecurGreenError) as e:
            plan.deposition(3, ipos, None, nseed=1)
        assert e.value.code == binding.RG_EINVAL
    wg = Waveguide(**p, batch_max=2)
    with pytest.raises(binding.RecurGreenError, match="Invalid final position index"):
        wg.compute_deposition(1, 0)
    wg.delete()
    plan.close()
