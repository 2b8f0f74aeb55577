"""Host-side logic of the autotuner (CPU): the probe guide keeps the slice spacing and the per-cell
disorder strength of the waveguide it stands for, and cache lookups are keyed by shape."""
import json
import math

from recurgreen_amd import autotune


def test_probe_params_keep_kdx_and_disorder_strength():
    p = dict(boundtype="dirichlet", nxdiso=1000, wol=64.25, aratio=50 / 64.25, dscat=5.0, ny=1285, dos_model=1)
    q = autotune.probe_params(p, 20)
    assert q["nxdiso"] == 20
    kdx = lambda d: 2 * math.pi * d["wol"] * d["aratio"] / d["nxdiso"]  # src/waveguide.f90:104,118
    assert math.isclose(kdx(q), kdx(p), rel_tol=1e-14)
    # sigma^2 is proportional to dscat / nxdiso (src/waveguide.f90:370)
    assert math.isclose(q["dscat"] / q["nxdiso"], p["dscat"] / p["nxdiso"], rel_tol=1e-14)
    assert q["ny"] == p["ny"] and q["wol"] == p["wol"]


def test_candidates_cover_every_precision_and_load_cache(tmp_path):
    assert set(autotune.CANDIDATES) == {"fp64", "mixed_fp32", "mixed_bf16x3"}
    for prec, opts in autotune.CANDIDATES.items():
        keys = [k for k, _ in opts]
        assert "project_parts" in keys
        assert ("inverse_fp64" in keys) == (prec == "fp64")
    path = tmp_path / "c.json"
    assert autotune.load_cache(str(path)) == {}
    path.write_text(json.dumps({"k": {"dyson": 1}}))
    assert autotune.load_cache(str(path)) == {"k": {"dyson": 1}}
