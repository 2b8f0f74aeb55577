"""LAPACK timing leg of the CPU baseline -- TEST INFRASTRUCTURE ONLY (imported by bench.py's cpu_baseline
and by tests/, never by the product).

The reference's FP64 path for a batch of realizations, on the host's LAPACK and BLAS (Intel MKL
through torch's CPU backend), one batched call per step of the reference's slice loop:
  - disorder: the oracle's Philox4x32-10 + Box-Muller generator (rgo_potential), scaled by stdu
    (src/waveguide.f90:366-372);
  - projection: V_mn = C(|m-n|) - C(m+n+2), C(q) = (1/ny) sum_i u_i cos(q pi y_i) (the oracle's transform
    form of modal^H diag(u) modal, src/waveguide.f90:377-380; Dirichlet guides);
  - slice matrices and Dyson step: Gss = inv(A_x - V_x - B Gss B) by zgetrf + zgetri (inverse_complex,
    src/base_utils.f90:235-274), Gs1 = -Gss (B Gs1) by zgemm (src/waveguide.f90:427-440);
  - Fisher-Lee and T = sigma(t)^2 (src/waveguide.f90:443-449); the singular values by MKL's gesdd, where
    the reference calls zgesvd (src/base_utils.f90:365): the same values to rounding.
Threads: torch's intra-op pool (MKL and the batched kernels).  tests/test_oracle.py pins this leg
against the oracle's own sweep.
"""
from __future__ import annotations

import concurrent.futures as cf

import numpy as np

from oracle import oracle as O


def tdistrib_lapack(w: O.Waveguide, seeds, threads: int = 1) -> np.ndarray:
    """Transmission eigenvalues (len(seeds), neigv), descending, of the oracle waveguide `w` for the given
    seeds.  Dirichlet guides only (the benchmark's)."""
    import torch

    if w.imin != 1 or w.jmin != 1 or w.imax != w.nmu or w.jmax != w.nmu:
        raise ValueError("the LAPACK leg takes the full aperture")
    torch.set_num_threads(max(1, int(threads)))
    n, ny, nxd, nx, nf = w.nmu, w.ny, w.nxdiso, w.nx, w.nxfree
    seeds = list(seeds)
    B = len(seeds)
    with cf.ThreadPoolExecutor(max(1, int(threads))) as ex:  # the generator releases the GIL
        U = np.stack(list(ex.map(w.potential, seeds)))  # (B, ny, nxdiso)
    U = torch.from_numpy(U)
    y = (np.arange(ny) + 0.5) / ny
    q = np.arange(2 * n + 2)
    cos_t = torch.from_numpy(np.cos(np.pi * np.outer(y, q)) / ny)  # (ny, 2n+2)
    m = np.arange(n)
    i1 = torch.from_numpy(np.abs(m[:, None] - m[None, :]))
    i2 = torch.from_numpy(m[:, None] + m[None, :] + 2)
    b = torch.from_numpy(w.b)
    bb = b[:, None] * b[None, :]
    diag_of = {"bound": torch.from_numpy(w.abound), "bulk": torch.from_numpy(w.abulk),
               "last": torch.from_numpy(w.abound + w.abulk if nx == 2 else w.abound)}

    def slice_matrix(ix):  # 1-based ix, src/waveguide.f90:357-380
        a = diag_of["bound"] if ix == 1 else (diag_of["last"] if ix == nx else diag_of["bulk"])
        M = torch.diag_embed(a).expand(B, n, n).clone()
        idd = ix - nf
        if 1 <= idd <= nxd:
            cq = U[:, :, idd - 1] @ cos_t  # (B, 2n+2)
            M -= (cq[:, i1] - cq[:, i2]).to(torch.complex128)
        return M

    gss = torch.linalg.inv(slice_matrix(1))
    gs1 = gss.clone()
    for ix in range(2, nx + 1):
        gss = torch.linalg.inv(slice_matrix(ix) - bb * gss)
        gs1 = -(gss @ (b[:, None] * gs1))
    sq = torch.from_numpy(np.sqrt(w.mu))
    t = 2j * (sq[:, None] * gs1 * sq[None, :])
    s = torch.linalg.svdvals(t)  # descending
    return (s * s).numpy()[:, :w.neigv]
