"""Host logic of the tprofile tasks (src/waveguide.f90:843-997, src/main.f90:466-483), no GPU."""
import numpy as np
import pytest

from recurgreen_amd import tprofile
from recurgreen_amd.binding import RecurGreenError


def test_window_and_group_sums_brute_force():
    rng = np.random.default_rng(2)
    nx, nv = 9, 7
    tv = np.sort(rng.uniform(0, 1, nv))[::-1]
    it = rng.uniform(0, 3, (nx, nv))
    prof, nt = tprofile.window_profile(tv, it, 0.3, 0.8)
    sel = [v for v in range(nv) if 0.3 <= tv[v] <= 0.8]
    assert nt == len(sel)
    ref = np.array([sum(it[x, v] for v in sel) for x in range(nx)]) * (4.0 / len(sel))
    np.testing.assert_allclose(prof, ref, rtol=1e-15)
    tm, dt = [0.2, 0.9, 0.5], [0.1, 0.05, 0.5]
    gp, gnt, tavg = tprofile.group_profiles(tv, it, tm, dt)
    for ip in range(3):
        s = [v for v in range(nv) if abs(tv[v] - tm[ip]) <= dt[ip]]
        assert gnt[ip] == len(s)
        r = np.array([sum(it[x, v] for v in s) for x in range(nx)])
        np.testing.assert_allclose(gp[:, ip], r * (4.0 / len(s)) if s else r, rtol=1e-15)
    assert tavg == pytest.approx(tv.sum() / nv, rel=1e-15)
    # compute_itx_profile's normalization: 2/(pi wol (tmax - tmin)), whatever nt
    wol = 7.3
    itx = tprofile.itx_window(tv, it, 0.3, 0.8, wol)
    ref_itx = np.array([sum(it[x, v] for v in sel) for x in range(nx)]) * 2.0 / (np.pi * wol * 0.5)
    np.testing.assert_allclose(itx, ref_itx, rtol=1e-15)
    np.testing.assert_allclose(itx * np.pi * wol * 0.5 / 2.0, prof * nt / 4.0, rtol=1e-14)
    assert not tprofile.itx_window(tv, it, 0.0, 1e-9, wol).any()
    # empty window: zero profile, nt = 0 (the realization is then left out of the average)
    p0, n0 = tprofile.window_profile(tv, it, 0.0, 1e-9)
    assert n0 == 0 and not p0.any()


def test_average_skips_empty_realizations():
    prof = np.array([[1.0, 2.0], [9.0, 9.0], [3.0, 4.0]])
    avg, ns = tprofile.average_profiles(prof, [2, 0, 1])
    assert ns == 2
    np.testing.assert_allclose(avg, [2.0, 3.0])


@pytest.mark.parametrize("tmin,tmax,msg", [(-0.1, 0.5, "Invalid Tmin"), (0.2, 1.5, "Invalid Tmax"),
                                           (0.5, 0.5, "Invalid range")])
def test_window_checks(tmin, tmax, msg):
    with pytest.raises(RecurGreenError, match=msg):
        tprofile.check_window(tmin, tmax)


def test_group_checks():
    with pytest.raises(RecurGreenError, match="Invalid transmission value"):
        tprofile.check_group([1.2], [0.1])
    with pytest.raises(RecurGreenError, match="half-width"):
        tprofile.check_group([0.5], [0.0])
    with pytest.raises(RecurGreenError, match="half-width"):
        tprofile.check_group([0.5], [0.6])
