"""The Fortran host (iso_c_binding module + namelist tdistrib driver) over the C ABI."""
import os
import subprocess

import numpy as np
import pytest

from recurgreen_amd import build as rgbuild
from recurgreen_amd.waveguide import generate_directory

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

NML = """&settings
    task = 'tdistrib'
/
&waveguide
    boundtype = 'dirichlet'
    wol = {wol}
    exfac = 1.0
    aratio = {aratio}
    dscat = 2.0
    dabso = 0.0
    nxdiso = 200
    nxfree = 0
    napera = 1.0
    naperb = 1.0
/
&tdistrib
    nseed = 16
    nbins = 64
    nthreads = 4
    save_eigvals = .true.
/
"""
GPU_GROUP = """&gpu
    device = 0
    precision = 'fp64'
    batch = 16
    ny = 205
    dos_model = 1
    bintype = 'chebyshev'
/
"""


def run(tmp_path, text):
    f = tmp_path / "settings.nml"
    f.write_text(text)
    return subprocess.run([rgbuild.FEXE, str(f)], cwd=tmp_path, capture_output=True, text=True, timeout=300)


def test_invalid_width_stops_like_the_reference(tmp_path):
    """Half-integer wol with Dirichlet boundaries stops before any device call
    (src/waveguide.f90:194-198)."""
    assert os.path.exists(rgbuild.FEXE), "Fortran host not built"
    r = run(tmp_path, NML.format(wol=10.5, aratio=1.0))
    # `stop "<message>"` exits with status 0, as the reference's `stop errmsg_invalid_arg` does
    assert "Invalid arguments" in r.stderr + r.stdout
    assert not (tmp_path / "out").exists()


def test_unknown_task_rejected(tmp_path):
    r = run(tmp_path, "&settings\n    task = 'nonsense'\n/\n")
    assert "Unknown task 'nonsense'" in r.stderr
    assert "Invalid arguments" in r.stderr


@pytest.mark.gpu
def test_fortran_tdistrib_matches_python_path(tmp_path):
    from oracle import oracle as O
    from recurgreen_amd import distrib

    print("[gpu-case] fortran host tdistrib", flush=True)
    r = run(tmp_path, NML.format(wol=10.25, aratio=10 / 10.25) + GPU_GROUP)
    assert r.returncode == 0 and "Result written" in r.stdout, r.stdout + r.stderr
    # the reference's result folder (generate_directory_waveguide, src/waveguide.f90:1046-1067)
    d = tmp_path / "out" / "dscat_2." / "dabso_0." / "wol_10.25" / "dirichlet" / "distrib"
    assert str(d) + "/" == str(tmp_path) + "/" + generate_directory(
        dict(dscat=2.0, dabso=0.0, wol=10.25, boundtype="dirichlet")) + "distrib/"
    eig = d / "result_1.eigvals.csv"
    dis = d / "result_1.distrib.csv"
    T = distrib.read_real_array(str(eig))
    ow = O.Waveguide("dirichlet", nxdiso=200, wol=10.25, aratio=10 / 10.25, dscat=2.0, ny=205, dos_model=1)
    To = ow.tdistrib_batch(1, 16, nthreads=4)
    np.testing.assert_allclose(T, To, rtol=1e-10, atol=1e-13)
    for path in (eig, dis):  # the %% header of save_distrib / save_raw_eigvals (src/main.f90:277-281,317-320)
        head = [ln for ln in path.read_text().splitlines() if ln.startswith("%%")]
        assert head[0].startswith("%% Computed on ") and " at " in head[0], head[0]
        assert any(ln.startswith("%% Nseed=16, Nbins=64, Nthreads=") and " seed/s" in ln for ln in head), head
    rho = distrib.load_distrib(str(dis))
    ref, miss = O.prepare_distrib(T, 64, "chebyshev")
    np.testing.assert_allclose(rho, ref, rtol=1e-12)


@pytest.mark.gpu
def test_fortran_tdistrib_refined_mixed(tmp_path):
    """precision = 'mixed_fp32' with refine = 2 in the &gpu group: the eigenvalues meet the FP64 bar
    against the oracle's (the plain complex64 sweep does not)."""
    from oracle import oracle as O
    from recurgreen_amd import distrib

    from test_gpu_parity import assert_T_close

    print("[gpu-case] fortran host tdistrib, mixed_fp32 refined", flush=True)
    grp = GPU_GROUP.replace("precision = 'fp64'", "precision = 'mixed_fp32'").replace("/\n", "    refine = 2\n/\n")
    r = run(tmp_path, NML.format(wol=10.25, aratio=10 / 10.25) + grp)
    assert r.returncode == 0 and "Result written" in r.stdout, r.stdout + r.stderr
    d = tmp_path / "out" / "dscat_2." / "dabso_0." / "wol_10.25" / "dirichlet" / "distrib"
    T = distrib.read_real_array(str(d / "result_1.eigvals.csv")).reshape(16, -1)
    ow = O.Waveguide("dirichlet", nxdiso=200, wol=10.25, aratio=10 / 10.25, dscat=2.0, ny=205, dos_model=1)
    To = ow.tdistrib_batch(1, 16, nthreads=4).reshape(16, -1)
    for k in range(16):
        assert_T_close(T[k], To[k])
