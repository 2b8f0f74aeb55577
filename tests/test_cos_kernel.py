"""The branch-free cosine of the disorder generator (recurgreen_amd/csrc/rg_cos.h), built for the host
with g++, against cos evaluated in 200-bit arithmetic (mpmath) and against the C library's cos that the
oracle's generator calls (oracle/rg_oracle.c: sqrt(-2 log u1) cos(2 pi u2)).

Points: 20 000 angles 2 pi u2 with u2 = b 2^-53 (the generator's), the 40 doubles on either side of
k pi/2 for k = 0..4 (where the reduced argument is smallest: 6.1e-17 next to pi/2), points around the
octant boundaries (k + 1/2) pi/2 where rint picks the quadrant, and tiny angles.  Bar: below 1 ulp of the
exact value everywhere (measured 0.72), and within 1 ulp of the C library's result (which is itself
within 0.56 ulp on these points)."""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "recurgreen_amd", "csrc", "rg_cos.h")

SRC = """#include <cmath>
#define RG_HD
#include "{header}"
extern "C" void cos_batch(const double *x, double *y, long n)
{{
    for (long i = 0; i < n; i++) y[i] = rg_cos_0_2pi(x[i]);
}}
"""


@pytest.fixture(scope="module")
def cos_lib(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("cos")
    (d / "c.cpp").write_text(SRC.format(header=os.path.abspath(HEADER)))
    so = d / "libc.so"
    subprocess.run(["g++", "-O2", "-Wno-unknown-pragmas", "-shared", "-fPIC", "-o", str(so), str(d / "c.cpp")],
                   check=True)
    return ctypes.CDLL(str(so))


def evaluate(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib.cos_batch(x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(len(x)))
    return y


def points():
    rng = np.random.default_rng(20261204)
    b = rng.integers(0, 2**53, 20000, dtype=np.int64)
    x = list((b.astype(np.float64) * 2.0**-53) * (2.0 * math.pi))
    for k in range(5):
        for direction in (0.0, 10.0):
            v = k * math.pi / 2
            for _ in range(40):
                x.append(v)
                v = math.nextafter(v, direction)
    for j in range(4):
        c = (j + 0.5) * math.pi / 2
        x += [c + d * 2e-16 for d in range(-20, 21)]
    x += [0.0, 5e-324, 1e-300, 1e-17, 1e-8, math.nextafter(2 * math.pi, 0)]
    return np.array([v for v in x if 0.0 <= v < 2 * math.pi])


def ulp_errors(x, y):
    mpmath.mp.prec = 200
    err = np.empty(len(x))
    for i, (xi, yi) in enumerate(zip(x, y)):
        exact = mpmath.cos(mpmath.mpf(float(xi)))
        err[i] = float(abs(mpmath.mpf(float(yi)) - exact) / math.ulp(float(exact)))
    return err


def test_cos_below_one_ulp(cos_lib):
    x = points()
    y = evaluate(cos_lib, x)
    err = ulp_errors(x, y)
    print(f"rg_cos_0_2pi: max error {err.max():.3f} ulp at x = {x[err.argmax()]!r}, mean {err.mean():.3f}")
    assert err.max() < 1.0


def test_cos_matches_the_c_library(cos_lib):
    x = points()
    y = evaluate(cos_lib, x)
    ref = np.array([math.cos(v) for v in x])
    ulps = np.array([math.ulp(r) for r in ref])
    assert np.max(np.abs(y - ref) / ulps) <= 1.0
    # the quadrant selection and signs (an error there is a whole unit, not an ulp)
    assert np.all(np.sign(y[np.abs(ref) > 1e-12]) == np.sign(ref[np.abs(ref) > 1e-12]))


def test_normal_cell_statistics(cos_lib):
    """Box-Muller with this cosine keeps the normal's moments (the generator's own check, on the host)."""
    rng = np.random.default_rng(7)
    n = 200000
    u1 = (rng.integers(0, 2**53, n, dtype=np.int64).astype(np.float64) + 1.0) * 2.0**-53
    u2 = rng.integers(0, 2**53, n, dtype=np.int64).astype(np.float64) * 2.0**-53
    z = np.sqrt(-2.0 * np.log(u1)) * evaluate(cos_lib, 2.0 * math.pi * u2)
    assert abs(z.mean()) < 5.0 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 8.0 / math.sqrt(n)
