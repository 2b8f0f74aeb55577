"""tpattern (src/main.f90:735-901) on the GPU: averaged channel probabilities from batched
complex128 reflection/transmission passes against the oracle's compute_refl_trans_v1 per seed.

Tolerances: probabilities agree with the oracle to 1e-10 relative to the largest probability;
flux conservation (zero absorption) holds to 1e-10; the Fortran host's CSV (16 significant
digits) matches the Python path to 1e-13 relative.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import build as rgbuild
from recurgreen_amd import tpattern
from recurgreen_amd.waveguide import Waveguide

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


CASES = {
    "dirichlet_N10_free": dict(boundtype="dirichlet", nxdiso=100, nxfree=20, wol=4.3, exfac=1.2, dscat=0.2),
    "periodic_N11": dict(boundtype="periodic", nxdiso=80, nxfree=5, wol=4.3, exfac=1.2, dscat=0.5),
    "N37_absorb": dict(boundtype="dirichlet", nxdiso=60, wol=18.7, dscat=3.0, dabso=0.5, ny=80),
    "periodic_N41_ny60": dict(boundtype="periodic", nxdiso=50, wol=20.3, dscat=2.0, ny=60, napera=0.5),
}


def oracle_tpattern(p, nseed, imode, probatype):
    ow = O.Waveguide(**p)
    imu = tpattern.mode_index(p["boundtype"], ow.nmu, imode)
    tsum = np.zeros(ow.nmu)
    rsum = np.zeros(ow.nmu)
    fpos, ipos = tpattern.deposition_positions(p)
    for s in range(1, nseed + 1):
        if probatype == "deposition":
            d = ow.deposition(ow.potential(s), fpos, ipos)
            tsum += np.abs(d[:, imu - 1]) ** 2
            continue
        r, t = ow.refl_trans(ow.potential(s), version=1)
        tp, rp = tpattern.channel_probabilities(r[None], t[None], probatype, imu)
        tsum += tp[0]
        rsum += rp[0]
    return tsum / nseed, rsum / nseed


@pytest.mark.parametrize("probatype", ["incident", "total", "deposition"])
@pytest.mark.parametrize("name", list(CASES))
def test_tpattern_matches_oracle(name, probatype):
    p = CASES[name]
    say(f"tpattern {name} {probatype}")
    wg = Waveguide(**p, batch_max=3)  # 7 seeds: passes of 3, 3, 1
    tp, rp, imu = tpattern.compute_tpattern(wg, 7, 2, probatype)
    to, ro = oracle_tpattern(p, 7, 2, probatype)
    scale = max(1.0, np.max(to), np.max(ro))
    assert np.max(np.abs(tp - to)) <= 1e-10 * scale
    assert np.max(np.abs(rp - ro)) <= 1e-10 * scale
    wg.delete()


def test_tpattern_flux_conservation_n128():
    """At zero absorption every column of the scattering matrix has unit norm
    (src/waveguide.test.f90:54-86): total tprob + rprob = 1 per incident channel, and the incident
    pattern sums to 1, on the N = 128 workload shape with a ragged last pass."""
    p = dict(boundtype="dirichlet", nxdiso=200, wol=64.25, aratio=10.0 / 64.25, dscat=1.0, ny=1285, dos_model=1)
    say("tpattern flux conservation N128")
    wg = Waveguide(**p, batch_max=16)
    tt, rt, _ = tpattern.compute_tpattern(wg, 40, 1, "total")
    assert np.max(np.abs(tt + rt - 1.0)) < 1e-10
    ti, ri, imu = tpattern.compute_tpattern(wg, 40, 1, "incident")
    assert imu == 1
    assert abs(ti.sum() + ri.sum() - 1.0) < 1e-10
    # the total pattern averages the same realizations: its frontal entry equals the incident sum
    assert abs(tt[0] - ti.sum()) < 1e-12 and abs(rt[0] - ri.sum()) < 1e-12
    wg.delete()


FORTRAN_NML = """&settings
    task = 'tpattern'
/
&waveguide
    boundtype = 'periodic'
    wol = 20.3
    exfac = 1.0
    aratio = 1.0
    dscat = 2.0
    dabso = 0.0
    nxdiso = 50
    nxfree = 0
    napera = 0.5
    naperb = 1.0
/
&tpattern
    nseed = 7
    imode = 3
    nthreads = 4
    probatype = '{probatype}'
/
&gpu
    device = 0
    precision = 'fp64'
    batch = 3
    ny = 60
    dos_model = 0
    bintype = 'chebyshev'
/
"""


@pytest.mark.parametrize("probatype", ["incident", "total", "deposition"])
def test_fortran_tpattern_matches_python_path(tmp_path, probatype):
    say(f"fortran host tpattern {probatype}")
    assert os.path.exists(rgbuild.FEXE), "Fortran host not built"
    f = tmp_path / "settings.nml"
    f.write_text(FORTRAN_NML.format(probatype=probatype))
    r = subprocess.run([rgbuild.FEXE, str(f)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Average profiles written" in r.stdout, r.stdout + r.stderr
    out = tmp_path / "out" / "dscat_2." / "dabso_0." / "wol_20.3" / "periodic" / "tpattern" / "result_1.csv"
    lines = out.read_text().splitlines()
    assert lines[4].startswith(f"%% Radiation pattern parameters: probatype='{probatype}', nseed=7, imode=3, time=")
    data = tpattern.load_tpattern(str(out))
    p = dict(boundtype="periodic", nxdiso=50, wol=20.3, dscat=2.0, ny=60, napera=0.5)
    wg = Waveguide(**p, batch_max=3)
    tp, rp, _ = tpattern.compute_tpattern(wg, 7, 3, probatype)
    wg.delete()
    np.testing.assert_array_equal(data[:, 0], np.arange(1, len(tp) + 1))
    np.testing.assert_allclose(data[:, 1], tp, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(data[:, 2], rp, rtol=1e-13, atol=1e-16)
