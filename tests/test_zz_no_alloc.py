"""rg_histogram_noalloc (no counterpart in the reference: the histogram of device-resident samples
with no allocation, its scratch a static device array of the library) against rg_histogram through
the C ABI, bit for bit: two calls in a row, the second with fewer samples and fewer bins, so that
counters left over from the first call would show; then the size limit and the argument checks."""
import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding, distrib

pytestmark = pytest.mark.gpu


def test_histogram_noalloc_matches_rg_histogram():
    import torch

    print("[gpu-case] rg_histogram_noalloc vs rg_histogram", flush=True)
    plan = binding.Plan(dict(boundtype="dirichlet", nxdiso=20, wol=4.3, dscat=0.5), batch_max=1)
    rng = np.random.default_rng(5)
    big = rng.uniform(-0.05, 1.05, 3_000_003)
    big[::1001] = np.nan
    big[7], big[8] = 1.0, 0.0
    small = rng.uniform(0.0, 1.0, 1001)
    small[3] = 2.0
    for samples, bt, nb in ((big, "chebyshev", 256), (small, "linear", 50)):
        e = distrib.bins(bt, nb)
        t = torch.from_numpy(samples).to("cuda:0")
        torch.cuda.synchronize()
        ref = plan.histogram(t, e)
        got = binding.histogram_noalloc(t, e, device=0)
        assert np.array_equal(got[0], ref[0]) and got[1:] == ref[1:], (bt, nb)
        h, miss = O.histogram(samples[~np.isnan(samples)], e)
        assert np.array_equal(got[0], h.astype(np.int64)) and got[1] == miss
        assert got[2] == int(np.isnan(samples).sum())
        print(f"[gpu-case] {bt} {nb}: {got[0].sum()} counted, {got[1]} missing, {got[2]} dropped", flush=True)
    e = distrib.bins("linear", 4097)
    with pytest.raises(binding.RecurGreenError) as err:
        binding.histogram_noalloc(t, e, device=0)
    assert err.value.code == binding.RG_ENOTSUP
    with pytest.raises(binding.RecurGreenError) as err:
        binding.histogram_noalloc(t, np.array([0.0, 0.5, 0.4, 1.0]), device=0)
    assert err.value.code == binding.RG_EINVAL
    plan.close()
