"""Host logic of the tpattern task (src/main.f90:735-901), no GPU calls."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding, tpattern
from recurgreen_amd import build as rgbuild
from recurgreen_amd.binding import RecurGreenError


@pytest.mark.parametrize("boundtype,wol", [("dirichlet", 4.3), ("dirichlet", 20.7), ("periodic", 4.3),
                                           ("periodic", 10.5), ("periodic", 25.98)])
def test_mode_index_is_frontal_mode(boundtype, wol):
    """imode=1 is the frontal mode (src/main.f90:766-772, settings.nml:102): the channel with the
    largest longitudinal wavenumber Re(mu) of the oracle's mode table."""
    ow = O.Waveguide(boundtype, nxdiso=10, wol=wol)
    imu = tpattern.mode_index(boundtype, ow.nmu, 1)
    assert imu - 1 == int(np.argmax(ow.mu.real)), (imu, ow.mu.real)


def test_mode_index_periodic_wraps():
    # nmu = 11: channels ordered by transverse index -5..5; imode 7..11 wrap to the start
    assert [tpattern.mode_index("periodic", 11, m) for m in range(1, 12)] == [6, 7, 8, 9, 10, 11, 1, 2, 3, 4, 5]
    assert [tpattern.mode_index("dirichlet", 11, m) for m in (1, 5, 11)] == [1, 5, 11]


def test_channel_probabilities_brute_force():
    rng = np.random.default_rng(3)
    nreal, n = 4, 7
    r = rng.normal(size=(nreal, n, n)) + 1j * rng.normal(size=(nreal, n, n))
    t = rng.normal(size=(nreal, n, n)) + 1j * rng.normal(size=(nreal, n, n))
    imu = 3
    tp, rp = tpattern.channel_probabilities(r, t, "incident", imu)
    tt, rt = tpattern.channel_probabilities(r, t, "total", imu)
    for k in range(nreal):
        for b in range(n):
            assert tp[k, b] == pytest.approx(abs(t[k, b, imu - 1]) ** 2, rel=1e-15)
            assert rp[k, b] == pytest.approx(abs(r[k, b, imu - 1]) ** 2, rel=1e-15)
        for a in range(n):  # column sums: total probability out of incident channel a
            assert tt[k, a] == pytest.approx(sum(abs(t[k, b, a]) ** 2 for b in range(n)), rel=1e-14)
            assert rt[k, a] == pytest.approx(sum(abs(r[k, b, a]) ** 2 for b in range(n)), rel=1e-14)


def test_argument_errors():
    with pytest.raises(RecurGreenError, match="Invalid number of disorder realizations"):
        tpattern.check_arguments(0, "incident")
    with pytest.raises(RecurGreenError, match="Invalid probatype"):
        tpattern.check_arguments(3, "reflected")
    tpattern.check_arguments(3, "deposition")
    assert tpattern.deposition_positions(dict(nxdiso=101, nxfree=7)) == (7 + 50 + 1, 1)  # src/main.f90:769-770


class _StubPlan:
    nmu, nprop, neigv, imin, imax, jmin, jmax = 5, 5, 5, 1, 5, 1, 5


class _StubGuide:
    params = dict(boundtype="dirichlet", wol=2.9, exfac=1.0, napera=1.0, naperb=1.0, dscat=0.5, dabso=0.0,
                  nxdiso=10, nxfree=0, aratio=1.0)
    plan = _StubPlan()


def test_save_load_round_trip(tmp_path):
    tp = np.array([0.1, 1.0 / 3.0, 2.5e-17, 0.0, 0.9999999999999999])
    rp = 1.0 - tp
    path = str(tmp_path / "tpattern" / "result_1.csv")
    tpattern.save_tpattern(path, _StubGuide(), tp, rp, "incident", 2, 10, 1.25)
    text = open(path).read().splitlines()
    assert text[0].startswith("%% Computed on ")
    assert text[4] == "%% Radiation pattern parameters: probatype='incident', nseed=10, imode=2, time=1.250 s."
    assert text[5] == "imu, tprob, rprob"
    data = tpattern.load_tpattern(path)
    np.testing.assert_array_equal(data[:, 0], np.arange(1, 6))
    # g0.16 keeps 16 significant digits (src/main.f90:889)
    np.testing.assert_allclose(data[:, 1], tp, rtol=1e-15, atol=0)
    np.testing.assert_allclose(data[:, 2], rp, rtol=1e-15, atol=0)


NML = """&settings
    task = 'tpattern'
/
&waveguide
    boundtype = 'periodic'
    wol = 4.3
    exfac = 1.0
    aratio = 1.0
    dscat = 0.5
    dabso = 0.0
    nxdiso = 50
    nxfree = 0
    napera = 1.0
    naperb = 1.0
/
&tpattern
    nseed = {nseed}
    imode = 1
    nthreads = 2
    probatype = '{probatype}'
/
"""


def _run(tmp_path, text):
    f = tmp_path / "settings.nml"
    f.write_text(text)
    return subprocess.run([rgbuild.FEXE, str(f)], cwd=tmp_path, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("nseed,probatype,msg", [
    (0, "incident", "Invalid number of disorder realizations (received 0)"),
    (4, "reflected", "Invalid probatype, received 'reflected'"),
])
def test_fortran_tpattern_argument_checks(tmp_path, nseed, probatype, msg):
    """The checks of src/main.f90:754-761 stop before any device call."""
    assert os.path.exists(rgbuild.FEXE), "Fortran host not built"
    r = _run(tmp_path, NML.format(nseed=nseed, probatype=probatype))
    assert msg in r.stderr, r.stderr
    assert "Invalid arguments" in r.stderr + r.stdout
    assert not (tmp_path / "out").exists()


class _OraclePlan:
    """Stand-in plan serving per-seed matrices from the oracle, to check the host's pass
    chunking and seed offsets without a device."""

    def __init__(self, ow, batch_max):
        self.ow, self.batch_max, self.nmu = ow, batch_max, ow.nmu
        self.calls = []

    def refl_trans(self, upot, seed0, nseed):
        self.calls.append((seed0, nseed))
        rt = [self.ow.refl_trans(self.ow.potential(s), version=1) for s in range(seed0, seed0 + nseed)]
        return np.stack([x[0] for x in rt]), np.stack([x[1] for x in rt]), np.zeros(nseed, np.int32), 0


@pytest.mark.parametrize("probatype", ["incident", "total"])
def test_compute_tpattern_chunking(probatype):
    p = dict(boundtype="periodic", nxdiso=30, wol=4.3, dscat=0.5)
    ow = O.Waveguide(**p)

    class G:
        params = p
        plan = _OraclePlan(ow, 3)

    tp, rp, imu = tpattern.compute_tpattern(G, 7, 2, probatype, seed0=4)
    assert G.plan.calls == [(4, 3), (7, 3), (10, 1)]
    tsum = np.zeros(ow.nmu)
    rsum = np.zeros(ow.nmu)
    for s in range(4, 11):
        r, t = ow.refl_trans(ow.potential(s), version=1)
        if probatype == "incident":
            tsum += np.abs(t[:, imu - 1]) ** 2
            rsum += np.abs(r[:, imu - 1]) ** 2
        else:
            tsum += (np.abs(t) ** 2).sum(axis=0)
            rsum += (np.abs(r) ** 2).sum(axis=0)
    np.testing.assert_allclose(tp, tsum / 7, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(rp, rsum / 7, rtol=1e-13, atol=1e-16)
    # zero absorption: unit flux out of every incident channel
    if probatype == "total":
        np.testing.assert_allclose(tp + rp, 1.0, rtol=0, atol=1e-11)
    else:
        assert abs(tp.sum() + rp.sum() - 1.0) < 1e-11
