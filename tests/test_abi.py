"""C ABI: the library loads and exports every symbol include/recurgreen_hip.h declares (CPU only,
no compute calls), and the host-side argument checks behave like the reference's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from recurgreen_amd import binding

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "recurgreen_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:rg_status|const char \*)\s*(rg_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert header_symbols() == sorted(binding.EXPORTS)


def test_library_exports_every_symbol():
    L = C.CDLL(binding.LIB_PATH)
    for name in header_symbols():
        assert hasattr(L, name), name
    assert binding.lib().rg_version().decode().startswith("recurgreen-hip")


def test_status_strings_keep_reference_messages():
    assert binding.status_string(binding.RG_EINVAL) == "RecurGreen: Invalid arguments, aborting..."
    assert "singular" in binding.status_string(binding.RG_ESINGULAR)


@pytest.mark.parametrize("kw", [dict(wol=4.5), dict(wol=0.0), dict(exfac=0.5), dict(dscat=-1.0),
                                dict(napera=0.0), dict(naperb=1.5), dict(nxfree=-1)])
def test_invalid_parameters_rejected_before_device(kw):
    """init_waveguide's checks (src/waveguide.f90:91-140,194-198) run before any device call."""
    params = dict(boundtype="dirichlet", nxdiso=50, wol=4.3)
    params.update(kw)
    with pytest.raises(binding.RecurGreenError) as e:
        binding.Plan(params, batch_max=1)
    assert e.value.code == binding.RG_EINVAL


def test_chunk_checks_of_compute_tdistrib():
    """src/waveguide.f90:407-421 run on the host before any device work."""
    from recurgreen_amd.waveguide import Waveguide

    wg = Waveguide.__new__(Waveguide)

    class FakePlan:
        neigv = 8

    wg.plan = FakePlan()
    with pytest.raises(binding.RecurGreenError, match="expected at least 8"):
        wg.compute_tdistrib(1, np.zeros(3))
    with pytest.raises(binding.RecurGreenError, match="expected 8"):
        wg.compute_tdistrib(1, np.zeros(12))
    with pytest.raises(binding.RecurGreenError, match=r"expected in \[1, 2\]"):
        wg.compute_tdistrib(3, np.zeros(16))


def test_potential_shape_checked_before_the_library_reads_it():
    """Every explicit-potential entry point takes upot(ix, iy[, r]) of shape (nxdiso, ny[, nreal]); any
    other shape is refused with RG_EINVAL on the host, before a pointer reaches the library."""
    from types import SimpleNamespace

    fake = SimpleNamespace(nxdiso=7, ny=5)
    u, n = binding.Plan._upot(fake, np.zeros((7, 5)))
    assert n == 1 and u.flags.f_contiguous and u.dtype == np.float64
    u, n = binding.Plan._upot(fake, np.zeros((7, 5, 3), dtype=np.float32))
    assert n == 3 and u.dtype == np.float64
    for bad in (np.zeros((5, 7)), np.zeros(35), np.zeros((7, 5, 0)), np.zeros((7, 5, 2, 1)), np.zeros((7, 6, 2))):
        with pytest.raises(binding.RecurGreenError) as e:
            binding.Plan._upot(fake, bad)
        assert e.value.code == binding.RG_EINVAL


def test_result_folder_follows_generate_directory():
    """The result folder of generate_directory_waveguide (src/waveguide.f90:1046-1067), whose
    docstring example is out/dscat_0.2/dabso_0./wol_25.5/periodic/: six decimals, a zero before the
    point below one, trailing zeros dropped (trim_zero, src/base_utils.f90:782-795)."""
    from recurgreen_amd.waveguide import generate_directory

    assert generate_directory(dict(dscat=0.2, dabso=0.0, wol=25.5, boundtype="periodic")) == \
        "out/dscat_0.2/dabso_0./wol_25.5/periodic/"
    assert generate_directory(dict(dscat=5.0, dabso=0.125, wol=64.25, boundtype="dirichlet"), "res") == \
        "res/dscat_5./dabso_0.125/wol_64.25/dirichlet/"
    assert generate_directory(dict(dscat=12.3456789, dabso=1e-7, wol=0.75, boundtype="dirichlet")) == \
        "out/dscat_12.345679/dabso_0./wol_0.75/dirichlet/"
