"""Full-size waveguides of the benchmark configurations, shared by the fixture generator
(tests/golden/make_fullsize.py) and the full-size parity tests (tests/test_gpu_fullsize.py).

Every guide is Dirichlet at wavelength 1 (k = 2 pi) with dx = dy = 0.05, white-noise disorder of
scattering length l_s (dscat = L / l_s), the continuum DOS 1/(4 pi) (dos_model = 1) and the
transverse grid ny = W / dy, i.e. the &waveguide mapping boundtype, wol = W, aratio = L / W,
nxdiso = L / dx of src/waveguide.f90:81 with the decoupled transverse grid.
"""
import numpy as np

DX = 0.05
K = 2.0 * np.pi


def guide(W, nx, ls, eps=0.0):
    """Dirichlet guide of width W, nx slices (L = nx dx), scattering length ls, absorption eps
    (dabso = eps L / k: mu = sqrt(1 + i dabso / kL - kperp^2), src/waveguide.f90:241)."""
    L = nx * DX
    return dict(boundtype="dirichlet", nxdiso=int(nx), nxfree=0, wol=float(W), exfac=1.0, aratio=L / W,
                dscat=L / ls, dabso=eps * L / K, napera=1.0, naperb=1.0, ny=int(round(W / DX)), dos_model=1)


def uneven_draws(nreal, seed=12345):
    """The uneven workload: N ~ U{16, 32, 64, 128, 256}, W = N/2 + 1/8, L / l_s ~ U[0.5, 10] with
    l_s = 10, eps ~ U[0, 0.05 k^2]; realization r (0-based) has seed r + 1."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nreal):
        N = int(rng.choice([16, 32, 64, 128, 256]))
        L = 10.0 * rng.uniform(0.5, 10.0)
        nx = max(2, int(round(L / DX)))
        eps = rng.uniform(0.0, 0.05 * K * K)
        out.append(guide(N / 2 + 0.125, nx, 10.0, eps))
    return out


def uneven_sample(per_n=2, nreal=10000, seed=12345):
    """The first `per_n` realizations of each width, in draw order, from the uneven workload of
    `nreal` realizations: (params list, seeds)."""
    draws = uneven_draws(nreal, seed)
    count, ps, seeds = {}, [], []
    for r, p in enumerate(draws):
        n = int(2 * p["wol"])
        if count.get(n, 0) < per_n:
            count[n] = count.get(n, 0) + 1
            ps.append(p)
            seeds.append(r + 1)
    return ps, seeds


# name: (params, seeds).  n128 is the headline workload (W = 64.25, L = 50, l_s = 10), c2 the FP64
# configuration (W = 50.25, L = 50, l_s = 10), c3 the N = 256 one (W = 128.25, l_s = 20) at its full
# length L = 100 and shortened to L = 10.
UNIFORM = {
    "n128": (guide(64.25, 1000, 10.0), [1, 2, 3, 4]),
    "c2": (guide(50.25, 1000, 10.0), [1, 2, 3, 4]),
    "c3_nx200": (guide(128.25, 200, 20.0), [1, 2]),
    "c3": (guide(128.25, 2000, 20.0), [1, 2]),
}
