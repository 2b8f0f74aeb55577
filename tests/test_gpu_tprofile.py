"""Transmission eigenstates inside the medium (rg_tprofile; compute_tprofile and
compute_tprofile_group, src/waveguide.f90:843-997) against the oracle: G(x;1) for every slice from
the reference's block-tridiagonal solve, t^H t decomposed by LAPACK (numpy eigh) on the host.

Tolerances (complex128): eigenvalues to 1e-10 relative (1e-13 absolute); per-eigenvector
intensities to 1e-8 of the slice's largest for eigenvalues separated by more than 1e-6 (an
eigenvector is defined up to rotations inside a degenerate eigenspace); window profiles, which do
not depend on that choice, to 1e-9 relative with window edges away from every eigenvalue."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding, tprofile
from recurgreen_amd.waveguide import Waveguide

pytestmark = pytest.mark.gpu

CASES = {
    "dirichlet_N10_free": (dict(boundtype="dirichlet", nxdiso=60, nxfree=4, wol=4.3, exfac=1.2, dscat=0.8), False),
    "periodic_N21_absorb": (dict(boundtype="periodic", nxdiso=40, wol=10.3, dscat=1.5, dabso=0.2), False),
    "N37_filtered_group": (dict(boundtype="dirichlet", nxdiso=30, wol=18.7, dscat=2.0, napera=0.6, naperb=0.9, ny=80),
                           True),
    "N64": (dict(boundtype="dirichlet", nxdiso=25, wol=32.25, aratio=2.0 / 32.25, dscat=1.0, ny=645, dos_model=1),
            False),
}


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def oracle_eigenstates(ow, u, filtered):
    """(T descending, intensity (nx, nv), U) from the oracle's Green column and numpy eigh."""
    G = ow.green_column(u, 1)
    s = np.sqrt(ow.mu.real)
    if filtered:
        i0, i1, j0, j1 = ow.imin - 1, ow.imax, ow.jmin - 1, ow.jmax
    else:
        i0, i1, j0, j1 = 0, ow.nmu, 0, ow.nmu
    t = 2j * s[i0:i1, None] * G[-1][i0:i1, j0:j1] * s[None, j0:j1]
    T, U = np.linalg.eigh(t.conj().T @ t)
    order = np.argsort(-T)
    T, U = T[order], U[:, order]
    W = s[j0:j1, None] * U
    inten = np.array([np.sum(np.abs(G[x][:, j0:j1] @ W) ** 2, axis=0) for x in range(ow.nx)])
    return T, inten


@pytest.mark.parametrize("name", list(CASES))
def test_tprofile_eigenstates_match_oracle(name):
    p, filtered = CASES[name]
    say(f"tprofile eigenstates {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=3)
    U3 = np.stack([ow.potential(s).T for s in (4, 5, 6)], axis=2)
    tv, it, st, code = plan.tprofile(U3, filtered=filtered)
    assert code == 0 and not st.any()
    for k in range(3):
        T, inten = oracle_eigenstates(ow, U3[:, :, k].T, filtered)
        err = np.abs(tv[k] - T)
        assert np.all(err <= np.maximum(1e-10 * np.abs(T), 1e-13)), err.max()
        gaps = np.abs(np.diff(T))
        iso = np.ones(T.size, bool)
        iso[:-1] &= gaps > 1e-6
        iso[1:] &= gaps > 1e-6
        scale = np.max(inten, axis=1, keepdims=True)
        assert np.max(np.abs(it[k][:, iso] - inten[:, iso]) / scale) <= 1e-8
        # window sums: edges between eigenvalues
        lo = 0.5 * (T[T.size // 2] + T[T.size // 2 + 1]) if T.size > 2 else 0.0
        prof, nt = tprofile.window_profile(tv[k], it[k], lo, 1.0)
        ref = inten[:, T >= lo].sum(axis=1) * 4.0 / max(1, int((T >= lo).sum()))
        assert nt == int((T >= lo).sum())
        np.testing.assert_allclose(prof, ref, rtol=1e-9, atol=1e-12 * ref.max())
        # compute_itx_profile: the same window, normalized by 2/(pi wol (tmax - tmin))
        itx = tprofile.itx_window(tv[k], it[k], lo, 1.0, p["wol"])
        sums = inten[:, T >= lo].sum(axis=1)
        np.testing.assert_allclose(itx * (np.pi * p["wol"] * (1.0 - lo) / 2.0), sums, rtol=1e-9,
                                   atol=1e-12 * max(sums.max(), 1.0))
    plan.close()


def test_tprofile_passes_and_host_calls():
    """Seeds over several device passes equal explicit potentials from the same generator, and the
    per-realization host calls (compute_tprofile / compute_tprofile_group signatures) equal the
    batch."""
    p = dict(boundtype="dirichlet", nxdiso=50, wol=20.25, aratio=0.5, dscat=1.5, ny=205, dos_model=1)
    say("tprofile seeded passes N40")
    wg = Waveguide(**p, batch_max=8)
    tv, it, st, code = wg.plan.tprofile(None, seed0=3, nseed=11, reserve=4)  # passes of 4, 4, 3
    assert code == 0 and not st.any()
    U = wg.plan.generate_potential(9, 2)
    tv2, it2, st2, code2 = wg.plan.tprofile(U)
    np.testing.assert_array_equal(tv2, tv[6:8])
    np.testing.assert_array_equal(it2, it[6:8])
    prof, nt = tprofile.compute_tprofile(wg, 5, 0.5, 1.0)
    ref, nref = tprofile.window_profile(tv[2], it[2], 0.5, 1.0)
    assert nt == nref
    np.testing.assert_array_equal(prof, ref)
    itx = wg.compute_itx_profile(5, 0.5, 1.0)
    np.testing.assert_array_equal(itx, tprofile.itx_window(tv[2], it[2], 0.5, 1.0, p["wol"]))
    gp, gnt, tavg = tprofile.compute_tprofile_group(wg, 5, [0.25, 0.75], [0.25, 0.25])
    assert tavg == pytest.approx(tv[2].sum() / tv.shape[1], rel=1e-14)
    assert gnt.sum() == int(((tv[2] >= 0.0) & (tv[2] <= 1.0)).sum()) + int((tv[2] == 0.5).sum())
    wg.delete()


def test_tprofile_requires_reserve_and_checks_apertures():
    p = dict(boundtype="dirichlet", nxdiso=20, wol=10.25, dscat=1.0, napera=1.0, naperb=0.5)
    say("tprofile argument checks")
    plan = binding.Plan(p, batch_max=2)
    tv = np.zeros((1, plan.nmu))
    it = np.zeros((1, plan.nx, plan.nmu))
    code = binding.lib().rg_tprofile(plan.handle, 1, 1, 0, binding.ptr(tv), binding.ptr(it), None)
    assert code == binding.RG_EINVAL  # no workspace yet
    with pytest.raises(binding.RecurGreenError) as e:  # more aperture columns than rows
        plan.tprofile(None, nseed=1, filtered=True)
    assert e.value.code == binding.RG_EINVAL
    plan.close()


def test_task_tprofile_files(tmp_path):
    """task_tprofile (src/main.f90:402-545): one raw row per realization, the average over the
    realizations with nt > 0 in 'x, tprofile' rows on the slice abscissas."""
    p = dict(boundtype="dirichlet", nxdiso=30, wol=10.25, dscat=2.0)
    say("task tprofile files")
    wg = Waveguide(**p, batch_max=4)
    raw, avgp, avg, nsample = tprofile.task_tprofile(wg, 6, 0.6, 1.0, outputdir=str(tmp_path), log=lambda m: None)
    profiles, nt = tprofile.compute_tprofile_batch(wg, 1, 6, 0.6, 1.0)
    assert nsample == int((nt > 0).sum()) and nsample > 0
    rows = [l for l in open(raw).read().splitlines() if not l.startswith("%")]
    assert len(rows) == 6
    got = np.array([[float(v) for v in r.split(",")] for r in rows])
    np.testing.assert_allclose(got, profiles, rtol=1e-15, atol=0)
    data = [l for l in open(avgp).read().splitlines() if not l.startswith("%")]
    assert data[0] == "x, tprofile"
    xa = np.array([[float(v) for v in r.split(",")] for r in data[1:]])
    np.testing.assert_allclose(xa[:, 0], wg.get_xmesh(), rtol=1e-15)
    np.testing.assert_allclose(xa[:, 1], profiles[nt > 0].mean(axis=0), rtol=1e-13)
    pr, n1 = wg.compute_tprofile(3, 0.6, 1.0)
    np.testing.assert_array_equal(pr, profiles[2])
    wg.delete()
