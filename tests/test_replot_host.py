"""Host logic of the PlotHistogram re-binning (src/plot_histogram.f90), no GPU calls: the device
count is replaced by the oracle's histogram, and the written file is compared with the reference's
fixture test/replot_1.distrib.csv (chebyshev, 256 bins, [0, 1] of test/test_data.csv)."""
import os
import shutil

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import distrib, replot
from recurgreen_amd.binding import RecurGreenError

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SETTINGS = """! settings of a re-binning run
&settings
    bintype = 'chebyshev'   ! lattice of the bins
    nbins = 256
    xmin = 0.0d0
    xmax = 1.0     ! upper end
/
"""


def test_parse_namelist():
    v = replot.parse_namelist(SETTINGS, "settings")
    assert v == {"bintype": "chebyshev", "nbins": 256, "xmin": 0.0, "xmax": 1.0}
    v = replot.parse_namelist("&settings bintype=\"linear\", nbins=8, xmin=-1.5E0, xmax=2.5 /\n", "settings")
    assert v == {"bintype": "linear", "nbins": 8, "xmin": -1.5, "xmax": 2.5}
    with pytest.raises(RecurGreenError):
        replot.parse_namelist("&other a=1 /", "settings")


@pytest.mark.parametrize("args,msg", [
    (("linear", 0, 0.0, 1.0), "Invalid number of bins (nbins=0 must be positive)"),
    (("linear", 4, 1.0, 1.0), "Invalid binning range"),
    (("linear", 4, float("nan"), 1.0), "Invalid binning range"),
    (("log", 4, 0.0, 1.0), "Unknown binning type 'log'"),
])
def test_settings_checks(args, msg):
    with pytest.raises(RecurGreenError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        replot.check_settings(*args)


def test_preamble_and_unique_name(tmp_path):
    shutil.copy(os.path.join(GOLD, "test_data.csv"), tmp_path / "data.csv")
    pre = replot.read_preamble(str(tmp_path / "data.csv"))
    assert len(pre) == 4 and all(p.startswith("%%") and p == p.rstrip() for p in pre)
    (tmp_path / "replot_1.distrib.csv").write_text("x")
    assert replot.unique_path(str(tmp_path), "replot_", ".distrib.csv").endswith("replot_2.distrib.csv")


def test_replot_reproduces_reference_fixture(tmp_path, monkeypatch):
    def oracle_counts(plan, samples, edges, device=0):
        h, miss = O.histogram(samples, edges)
        return np.asarray(h), miss

    monkeypatch.setattr(replot, "device_counts", oracle_counts)
    shutil.copy(os.path.join(GOLD, "test_data.csv"), tmp_path / "test_data.csv")
    (tmp_path / "plothisto.nml").write_text(SETTINGS)
    logs = []
    out = replot.replot(str(tmp_path / "plothisto.nml"), str(tmp_path / "test_data.csv"), plan=object(),
                        log=logs.append)
    assert os.path.basename(out) == "replot_1.distrib.csv"
    assert any("6 missing samples" in l for l in logs)
    got = open(out).read().splitlines()
    ref = open(os.path.join(GOLD, "replot_1.distrib.csv")).read().splitlines()
    assert got[:5] == [l.rstrip() for l in ref[:5]]  # preamble and header
    np.testing.assert_allclose(distrib.load_distrib(out), distrib.load_distrib(os.path.join(GOLD, "replot_1.distrib.csv")),
                               rtol=1e-15, atol=0)


def test_histogram_plan_arguments_pass_the_host_checks():
    """The minimal plan's parameters pass init_waveguide's argument checks; without a device the
    create call then stops at the device query (RG_EDEVICE), not at the arguments (RG_EINVAL)."""
    import torch

    from recurgreen_amd import binding

    if torch.cuda.is_available():
        pytest.skip("a device is present: the GPU tests create the plan")
    with pytest.raises(RecurGreenError) as e:
        replot.histogram_plan(0)
    assert e.value.code == binding.RG_EDEVICE
