"""PlotHistogram re-binning (src/plot_histogram.f90) with the device histogram: the reference's
fixture test/replot_1.distrib.csv reproduced from test/test_data.csv (rows to 1e-15 relative,
preamble and header verbatim), and counts of a large sample bit-exact against the oracle."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import distrib, replot

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SETTINGS = "&settings\n    bintype = 'chebyshev'\n    nbins = 256\n    xmin = 0.0\n    xmax = 1.0\n/\n"


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_replot_fixture_on_device(tmp_path):
    say("replot fixture")
    shutil.copy(os.path.join(GOLD, "test_data.csv"), tmp_path / "test_data.csv")
    (tmp_path / "plothisto.nml").write_text(SETTINGS)
    r = subprocess.run([sys.executable, "-m", "recurgreen_amd.replot", str(tmp_path / "plothisto.nml"),
                        str(tmp_path / "test_data.csv")], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "6 missing samples" in r.stdout
    out = tmp_path / "replot_1.distrib.csv"
    ref = os.path.join(GOLD, "replot_1.distrib.csv")
    assert out.read_text().splitlines()[:5] == [l.rstrip() for l in open(ref).read().splitlines()[:5]]
    np.testing.assert_allclose(distrib.load_distrib(str(out)), distrib.load_distrib(ref), rtol=1e-15, atol=0)


@pytest.mark.parametrize("bintype,xmin,xmax", [("linear", -0.25, 1.0), ("chebyshev", 0.0, 1.0)])
def test_replot_counts_match_oracle(bintype, xmin, xmax):
    say(f"replot counts {bintype}")
    rng = np.random.default_rng(11)
    x = rng.uniform(-0.5, 1.2, 3_000_000)
    x[::1000] = np.nan
    x[5] = 1.0
    edges = distrib.bins(bintype, 300, xmin, xmax)
    plan = replot.histogram_plan(0)
    counts, missing = replot.device_counts(plan, x, edges, 0)
    plan.close()
    ho, mo = O.histogram(x, edges)
    np.testing.assert_array_equal(counts, ho)
    assert missing == mo
