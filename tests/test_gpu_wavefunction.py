"""compute_wavefunction (src/waveguide.f90:636-697) through the C ABI against the oracle: the column
G(x; ipos)[:, imu] of the reference's block-tridiagonal solve for every slice, through the oracle's
modal matrix.  Tolerance (complex128): 1e-10 of max|psi|."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding, tpattern
from recurgreen_amd.waveguide import Waveguide

pytestmark = pytest.mark.gpu

CASES = {
    "dirichlet_N10_free": dict(boundtype="dirichlet", nxdiso=40, nxfree=3, wol=4.3, dscat=0.8),
    "periodic_N21_absorb": dict(boundtype="periodic", nxdiso=30, wol=10.3, dscat=1.5, dabso=0.2),
    "dirichlet_N20_ny205": dict(boundtype="dirichlet", nxdiso=30, wol=10.25, aratio=0.5, dscat=2.0, ny=205, dos_model=1),
    "periodic_N41_ny60": dict(boundtype="periodic", nxdiso=20, nxfree=2, wol=20.3, dscat=1.0, ny=60),
}


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", list(CASES))
def test_wavefunction_matches_oracle(name):
    p = CASES[name]
    say(f"wavefunction {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=1)
    u = ow.potential(7)
    for ipos, imode in [(1, 1), (1, 3), (ow.nx // 2, 2), (ow.nx, 1)]:
        psi, st, code = plan.wavefunction(ipos, imode, u.T)
        assert code == 0 and not st.any()
        imu = tpattern.mode_index(p["boundtype"], ow.nmu, imode)
        G = ow.green_column(u, ipos)
        ref = np.array([ow.modal @ G[x][:, imu - 1] for x in range(ow.nx)])
        assert np.max(np.abs(psi - ref)) <= 1e-10 * np.max(np.abs(ref)), (ipos, imode)
    plan.close()


def test_wfundata_layout_and_checks():
    p = CASES["dirichlet_N10_free"]
    say("wavefunction wfundata")
    wg = Waveguide(**p, batch_max=1)
    w = wg.compute_wavefunction(5, 1, 2)
    nx, ny = wg.plan.nx, wg.plan.ny
    assert w.shape == (nx * ny, 5)
    psi, st, code = wg.plan.wavefunction(1, 2, None, seed=5)
    np.testing.assert_array_equal(w[:, 2].reshape(nx, ny), psi.real)
    u = wg.generate_potential(5, 1)[:, :, 0]
    np.testing.assert_array_equal(w[3 * ny:(3 + p["nxdiso"]) * ny, 4], u.ravel())
    assert not w[:3 * ny, 4].any() and not w[(3 + p["nxdiso"]) * ny:, 4].any()
    np.testing.assert_allclose(w[::ny, 0], wg.get_xmesh())
    with pytest.raises(binding.RecurGreenError, match="Invalid initial mode index"):
        wg.compute_wavefunction(5, 1, wg.plan.nmu + 1)
    with pytest.raises(binding.RecurGreenError, match="Invalid initial position index"):
        wg.compute_wavefunction(5, 0, 1)
    wg.delete()
