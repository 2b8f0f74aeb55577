"""Uneven batches (rg_tdistrib_uneven): realizations with their own length, absorption, disorder
strength, seed and width (up to N = 256) in one call.  Each realization must equal, bit for bit, the uniform
one-realization sweep of the same parameters and seed (same kernels, same data, constants taken
per realization instead of shared); rejected parameters and unsupported widths are reported per
realization.  One absorbing realization is also checked against the FP64 oracle (1e-10 relative).
"""
import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding

from test_gpu_parity import assert_T_close, say

pytestmark = pytest.mark.gpu

DX, LS = 0.05, 5.0


def guide(wol, nxdiso, dabso=0.0, ny=None, nxfree=0, exfac=1.0, dos_model=1):
    """Dirichlet guide of width wol at dx = 0.05 and l_s = 5 (dscat = L / l_s)."""
    L = nxdiso * DX
    return dict(boundtype="dirichlet", wol=wol, exfac=exfac, nxdiso=nxdiso, nxfree=nxfree, aratio=L / wol,
                dscat=L / LS, dabso=dabso, ny=ny, dos_model=dos_model)


def uneven_set():
    ps = [
        guide(10.25, 30, 0.0, ny=205), guide(10.25, 55, 0.3, ny=205), guide(18.7, 20, 0.5, ny=80),
        guide(10.25, 55, 0.0, ny=205), guide(10.25, 2, 0.0, ny=205),   # nx = 2: last diagonal = bound + bulk
        guide(4.3, 10, 0.2, exfac=1.2, nxfree=5, dos_model=0), guide(10.25, 80, 1.0, ny=205),
        guide(10.0, 20),                                            # wol a multiple of 0.5: rejected
        guide(150.25, 10, ny=301),                                  # N = 300 > 256: not supported
        guide(18.7, 40, 0.0, ny=80), guide(4.3, 25, 0.0, exfac=1.2, nxfree=5, dos_model=0),
        guide(10.25, 12, 0.1, ny=205), guide(100.25, 6, 0.2, ny=201), guide(100.25, 9, 0.0, ny=201),  # N = 200
    ]
    seeds = np.arange(100, 100 + len(ps), dtype=np.int64)
    return ps, seeds


@pytest.mark.parametrize("precision", ["fp64", "mixed_fp32"])
def test_uneven_batch_equals_uniform_sweeps(precision):
    ps, seeds = uneven_set()
    say(f"uneven batch of {len(ps)} ({precision})")
    vals, st, code = binding.tdistrib_uneven(ps, seeds, precision, batch_max=4)
    assert st[7] == binding.RG_EINVAL and vals[7].size == 0
    assert st[8] == binding.RG_ENOTSUP
    assert code != 0
    for r, p in enumerate(ps):
        if r in (7, 8):
            continue
        assert st[r] == 0, (r, st[r])
        plan = binding.Plan(p, batch_max=1)
        T, s1, c1 = plan.tdistrib(int(seeds[r]), 1, precision)
        plan.close()
        assert c1 == 0 and vals[r].shape == T.shape, r
        assert np.array_equal(vals[r], T), (r, np.max(np.abs(vals[r] - T)))


def test_uneven_absorbing_realization_against_oracle():
    ps, seeds = uneven_set()
    say("uneven batch vs oracle")
    vals, st, _ = binding.tdistrib_uneven(ps, seeds, "fp64", batch_max=8)
    for r in (1, 6):  # dabso 0.3 and 1.0, different lengths in one group
        ow = O.Waveguide(**ps[r])
        assert st[r] == 0
        assert_T_close(vals[r], ow.tdistrib_potential(ow.potential(int(seeds[r]))))


@pytest.mark.parametrize("precision", ["fp64", "mixed_fp32"])
def test_split_group_equals_unsplit(precision):
    """A group of 512 or more realizations above 128 modes is swept as two halves on two streams; each
    realization must equal, bit for bit, the same realization in a call whose groups are not split
    (the same realizations in two calls of 260)."""
    rng = np.random.default_rng(7)
    ps = [guide(128.25, int(rng.integers(2, 7)), float(rng.uniform(0.0, 0.5)), ny=257) for _ in range(520)]
    seeds = np.arange(1000, 1520, dtype=np.int64)
    say(f"split N=256 group of 520 ({precision})")
    vals, st, code = binding.tdistrib_uneven(ps, seeds, precision, batch_max=520)
    assert code == 0 and not st.any()
    for lo in (0, 260):
        ref, st2, code2 = binding.tdistrib_uneven(ps[lo:lo + 260], seeds[lo:lo + 260], precision, batch_max=260)
        assert code2 == 0 and not st2.any()
        for k in range(260):
            assert np.array_equal(vals[lo + k], ref[k]), lo + k
