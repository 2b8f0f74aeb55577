"""The Fortran host's wavefunction, tprofile and tprofile_group tasks (src/main.f90:335-735) against
the Python path over the same C ABI; their argument checks stop before any device call."""
import os
import subprocess

import numpy as np
import pytest

from recurgreen_amd import build as rgbuild
from recurgreen_amd import tprofile
from recurgreen_amd.waveguide import Waveguide, generate_directory

WG = """&waveguide
    boundtype = '{bt}'
    wol = {wol}
    exfac = 1.0
    aratio = 1.0
    dscat = 1.5
    dabso = 0.0
    nxdiso = 30
    nxfree = 2
    napera = {na}
    naperb = 1.0
/
&gpu
    device = 0
    batch = 4
/
"""
PARAMS = dict(nxdiso=30, nxfree=2, dscat=1.5)


def out_dir(bt, wol):
    """The reference's result folder for the guides of WG (src/waveguide.f90:1046-1067)."""
    return generate_directory(dict(dscat=1.5, dabso=0.0, wol=wol, boundtype=bt))


def run(tmp_path, text):
    f = tmp_path / "settings.nml"
    f.write_text(text)
    return subprocess.run([rgbuild.FEXE, str(f)], cwd=tmp_path, capture_output=True, text=True, timeout=300)


def rows(path, skip_header=True):
    out = []
    for line in open(path).read().splitlines():
        if line.startswith("%") or (skip_header and line[:1].isalpha()):
            continue
        out.append([float(v) for v in line.split(",")])
    return np.array(out)


@pytest.mark.parametrize("task,group,msg", [
    ("tprofile", "&tprofile\n    tmin = 0.9\n    tmax = 0.5\n    nseed = 4\n    nthreads = 1\n/\n", "Invalid range (Tmin >= Tmax)"),
    ("tprofile", "&tprofile\n    tmin = 0.1\n    tmax = 0.5\n    nseed = 0\n    nthreads = 1\n/\n", "Invalid number of disorder realizations"),
    ("tprofile_group", "&tprofile_group\n    tm = 0.1, 0.2, 0.3, 0.4\n    dt = 0.1, 0.1, 0.6, 0.1\n    nseed = 4\n    nthreads = 1\n/\n",
     "Invalid transmission half-width"),
])
def test_task_argument_checks(tmp_path, task, group, msg):
    assert os.path.exists(rgbuild.FEXE), "Fortran host not built"
    r = run(tmp_path, f"&settings\n    task = '{task}'\n/\n" + WG.format(bt="dirichlet", wol=4.3, na=1.0) + group)
    assert msg in r.stderr, r.stderr
    assert not (tmp_path / "out").exists()


@pytest.mark.gpu
def test_fortran_wavefunction_matches_python(tmp_path):
    print("[gpu-case] fortran wavefunction", flush=True)
    text = "&settings\n    task = 'wavefunction'\n/\n" + WG.format(bt="periodic", wol=4.3, na=1.0) + \
        "&wavefunction\n    iseed = 3\n    ipos = 5\n    imode = 2\n/\n"
    r = run(tmp_path, text)
    assert r.returncode == 0 and "Result written" in r.stdout, r.stdout + r.stderr
    data = rows(tmp_path / out_dir("periodic", 4.3) / "wavefun" / "result_1.wavefun.csv")
    wg = Waveguide("periodic", wol=4.3, **PARAMS, batch_max=1)
    ref = wg.compute_wavefunction(3, 5, 2)
    wg.delete()
    np.testing.assert_allclose(data, ref, rtol=1e-13, atol=1e-15 * np.abs(ref).max())


@pytest.mark.gpu
def test_fortran_tprofile_matches_python(tmp_path):
    print("[gpu-case] fortran tprofile", flush=True)
    text = "&settings\n    task = 'tprofile'\n/\n" + WG.format(bt="dirichlet", wol=10.25, na=1.0) + \
        "&tprofile\n    tmin = 0.5\n    tmax = 1.0\n    nseed = 9\n    nthreads = 4\n/\n"
    r = run(tmp_path, text)
    assert r.returncode == 0 and "Average profile written" in r.stdout, r.stdout + r.stderr
    raw = rows(tmp_path / out_dir("dirichlet", 10.25) / "tprofile" / "result_1.raw.csv")
    avg = rows(tmp_path / out_dir("dirichlet", 10.25) / "tprofile" / "result_1.avg.csv")
    wg = Waveguide("dirichlet", wol=10.25, **PARAMS, batch_max=4)
    prof, nt = tprofile.compute_tprofile_batch(wg, 1, 9, 0.5, 1.0)
    ref_avg, ns = tprofile.average_profiles(prof, nt)
    np.testing.assert_allclose(raw, prof, rtol=1e-13, atol=1e-15 * prof.max())
    np.testing.assert_allclose(avg[:, 0], wg.get_xmesh(), rtol=1e-15)
    np.testing.assert_allclose(avg[:, 1], ref_avg, rtol=1e-12)
    wg.delete()


@pytest.mark.gpu
def test_fortran_tprofile_group_matches_python(tmp_path):
    print("[gpu-case] fortran tprofile_group", flush=True)
    tm, dt = [0.1, 0.4, 0.7, 0.95], [0.1, 0.1, 0.1, 0.05]
    text = "&settings\n    task = 'tprofile_group'\n/\n" + WG.format(bt="dirichlet", wol=10.25, na=0.7) + \
        "&tprofile_group\n    tm = 0.1, 0.4, 0.7, 0.95\n    dt = 0.1, 0.1, 0.1, 0.05\n    nseed = 10\n    nthreads = 4\n/\n"
    r = run(tmp_path, text)
    assert r.returncode == 0 and "Average profiles written" in r.stdout, r.stdout + r.stderr
    path = tmp_path / out_dir("dirichlet", 10.25) / "tprofile_group" / "result_1.csv"
    text_lines = open(path).read().splitlines()
    assert text_lines[6].startswith("x, tprofile") and text_lines[6].count("tprofile") == 4
    data = rows(path)
    wg = Waveguide("dirichlet", wol=10.25, napera=0.7, **PARAMS, batch_max=4)
    prof, nt, tavg = tprofile.compute_tprofile_group_batch(wg, 1, 10, tm, dt)
    for ip in range(4):
        ref, ns = tprofile.average_profiles(prof[:, :, ip], nt[:, ip])
        if ns:
            np.testing.assert_allclose(data[:, 1 + ip], ref, rtol=1e-12)
    assert "over nseed=10, tavg=" in text_lines[5]
    tl = text_lines[5].split("tavg=")[1].split("+-")[0]
    assert float(tl) == pytest.approx(float(np.mean(tavg)), rel=1e-5)
    wg.delete()
