"""Host-side mirror of RecurGreen's `waveguide` module over the gfx950 library.

`Waveguide` keeps the names, argument meaning and error behaviour of waveguide_t
(src/waveguide.f90:13-73): init (:81), get_nmu/get_neigv/get_nx (:291-314),
compute_tdistrib (:392), del (:275).  The reference computes one realization per call
(chunk `iseed` of `tallvals`) inside an OpenMP loop (src/main.f90:148-171); here
compute_tdistrib_batch runs the whole loop as one device pass, and compute_tdistrib keeps the
single-chunk form for callers written against the reference.
"""
from __future__ import annotations

import numpy as np

from . import binding
from .binding import RecurGreenError

ERRMSG_INVALID_ARG = "RecurGreen: Invalid arguments, aborting..."  # src/base_utils.f90:56


def _folder_value(name: str, x: float) -> str:
    """name + x with six decimals, a zero before the point below one and the trailing zeros
    dropped (trim_zero, src/base_utils.f90:782-795): 0.2 -> "0.2", 5 -> "5.", 0 -> "0."."""
    return name + f"{x:.6f}".rstrip("0")


def generate_directory(params: dict, root: str = "out") -> str:
    """Result folder of a waveguide with these parameters (src/waveguide.f90:1046-1067)."""
    return "/".join([root, _folder_value("dscat_", params["dscat"]), _folder_value("dabso_", params["dabso"]),
                     _folder_value("wol_", params["wol"]), params["boundtype"]]) + "/"


class Waveguide:
    def __init__(self, boundtype: str = "dirichlet", nxdiso: int = 100, nxfree: int = 0, wol: float = 4.3,
                 exfac: float = 1.0, aratio: float = 1.0, dscat: float = 0.2, dabso: float = 0.0,
                 napera: float = 1.0, naperb: float = 1.0, *, ny: int | None = None, dos_model: int = 0,
                 batch_max: int = 256, device: int = 0):
        self.params = dict(boundtype=boundtype, nxdiso=nxdiso, nxfree=nxfree, wol=wol, exfac=exfac,
                           aratio=aratio, dscat=dscat, dabso=dabso, napera=napera, naperb=naperb,
                           ny=ny, dos_model=dos_model)
        if boundtype not in binding.BOUNDTYPES:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Unknown boundary condition type '{boundtype}', aborting...")
        self.plan = binding.Plan(self.params, batch_max=batch_max, device=device)

    # -- accessors (src/waveguide.f90:291-314) ---------------------------------------------
    def get_nmu(self) -> int:
        return self.plan.nmu

    def get_neigv(self) -> int:
        return self.plan.neigv

    def get_nx(self) -> int:
        return self.plan.nx

    def get_xmesh(self) -> np.ndarray:
        """x_i/L (src/waveguide.f90:146)."""
        p = self.params
        i = np.arange(1, self.plan.nx + 1)
        return p["aratio"] * (i - 0.5 - p["nxfree"]) / p["nxdiso"]

    def get_parameters(self) -> dict:
        return dict(self.params)

    def generate_directory(self, root: str = "out") -> str:
        """generate_directory_waveguide (src/waveguide.f90:1046-1067): the result folder
        <root>/dscat_<L/ls>/dabso_<L/la>/wol_<W/lambda>/<boundtype>/."""
        return generate_directory(self.params, root)

    # -- transmission eigenvalues -------------------------------------------------------------
    def compute_tdistrib(self, iseed: int, tallvals: np.ndarray, precision: str = "fp64") -> None:
        """Chunk `iseed` (1-based) of tallvals <- transmission eigenvalues of realization iseed,
        with the argument checks of src/waveguide.f90:407-421."""
        neigv = self.plan.neigv
        nchunk = tallvals.size // neigv
        if nchunk == 0:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Invalid eigenvalue array length, received {tallvals.size}, expected at least {neigv}...")
        if nchunk * neigv != tallvals.size:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Invalid eigenvalue array length, received {tallvals.size}, expected {nchunk * neigv}...")
        if iseed < 1 or nchunk < iseed:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Invalid chunk index, received {iseed}, expected in [1, {nchunk}]...")
        T, st, code = self.plan.tdistrib(iseed, 1, precision)
        if code:
            raise RecurGreenError(code)
        tallvals[(iseed - 1) * neigv: iseed * neigv] = T

    def compute_tdistrib_batch(self, seed0: int, nseed: int, precision: str = "fp64"):
        """All realizations seed0 .. seed0+nseed-1 in one device pass (the OpenMP loop of
        src/main.f90:153-166).  Returns (tallvals, per-realization status)."""
        T, st, code = self.plan.tdistrib(seed0, nseed, precision)
        if code and code not in (binding.RG_ESINGULAR, binding.RG_ENOCONV, binding.RG_ENONFINITE):
            raise RecurGreenError(code)
        return T, st

    def compute_tdistrib_potential(self, upot: np.ndarray, precision: str = "fp64"):
        """Explicit potentials upot(ix, iy[, r]) in U*dx/k units (src/waveguide.f90:43 layout)."""
        T, st, code = self.plan.tdistrib_potential(upot, precision)
        return T, st

    def transmission_matrix(self, upot: np.ndarray | None, precision: str = "fp64") -> np.ndarray:
        t, st, code = self.plan.transmission_potential(upot, precision=precision)
        if code:
            raise RecurGreenError(code)
        return t[0] if t.shape[0] == 1 else t

    def compute_refl_trans(self, iseed: int):
        """(rmat, tmat) of realization `iseed`, as compute_refl_trans_v1 (src/waveguide.f90:469-519)
        returns them: nmu x nmu complex128, with the sqrt(Re mu) Fisher-Lee factors of :510-512."""
        r, t, st, code = self.plan.refl_trans(None, seed0=iseed, nseed=1)
        if code:
            raise RecurGreenError(code)
        return r[0], t[0]

    def compute_deposition(self, iseed: int, fpos: int, ipos: int = 1) -> np.ndarray:
        """depmat of realization `iseed` as compute_deposition (src/waveguide.f90:710-745) returns it:
        nmu x nmu complex128, 2i s G(fpos; ipos) s with s = sqrt(Re mu); same position checks
        (:719-727)."""
        nx = self.plan.nx
        if fpos < 1 or nx < fpos:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Invalid final position index, received fpos={fpos}, expected in 1..{nx}, aborting...")
        if ipos < 1 or nx < ipos:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Invalid initial position index, received ipos={ipos}, expected in 1..{nx}, aborting...")
        d, st, code = self.plan.deposition(fpos, ipos, None, seed0=iseed, nseed=1)
        if code:
            raise RecurGreenError(code)
        return d[0]

    def compute_wavefunction(self, iseed: int, ipos: int, imode: int) -> np.ndarray:
        """wfundata of compute_wavefunction (src/waveguide.f90:636-697): (nx*ny, 5) rows
        [x/L, y/W, Re psi, Im psi, Upot] with row (ix-1)*ny + iy, Upot = U dx/k on the disordered
        slices and 0 on the free ones; same index checks (:646-654)."""
        nmu, nx, ny = self.plan.nmu, self.plan.nx, self.plan.ny
        if imode < 1 or nmu < imode:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Invalid initial mode index, received imode={imode}, expected in 1..{nmu}, aborting...")
        if ipos < 1 or nx < ipos:
            raise RecurGreenError(binding.RG_EINVAL,
                                  f"Invalid initial position index, received ipos={ipos}, expected in 1..{nx}, aborting...")
        psi, st, code = self.plan.wavefunction(ipos, imode, None, seed=iseed)
        if code:
            raise RecurGreenError(code)
        w = np.zeros((nx * ny, 5))
        w[:, 0] = np.repeat(self.get_xmesh(), ny)
        w[:, 1] = np.tile((np.arange(1, ny + 1) - 0.5) / ny, nx)
        w[:, 2] = psi.real.ravel()
        w[:, 3] = psi.imag.ravel()
        u = self.plan.generate_potential(iseed, 1)[:, :, 0]  # (nxdiso, ny)
        nf = self.params.get("nxfree", 0)
        w[nf * ny:(nf + u.shape[0]) * ny, 4] = u.ravel()
        return w

    def compute_tprofile(self, iseed: int, tmin: float, tmax: float):
        """compute_tprofile (src/waveguide.f90:843-907): (tprofile[nx], nt) of realization iseed."""
        from . import tprofile

        return tprofile.compute_tprofile(self, iseed, tmin, tmax)

    def compute_itx_profile(self, iseed: int, tmin: float, tmax: float):
        """compute_itx_profile (src/waveguide.f90:764-825): itx[nx] of realization iseed."""
        from . import tprofile

        return tprofile.compute_itx_profile(self, iseed, tmin, tmax)

    def compute_tprofile_group(self, iseed: int, tm, dt):
        """compute_tprofile_group (src/waveguide.f90:924-997): (tprofile[nx, nprofile], nt, tavg)."""
        from . import tprofile

        return tprofile.compute_tprofile_group(self, iseed, tm, dt)

    def generate_potential(self, seed0: int, nreal: int = 1) -> np.ndarray:
        return self.plan.generate_potential(seed0, nreal)

    def delete(self) -> None:
        """waveguide_t%del (src/waveguide.f90:275)."""
        self.plan.close()

    del_ = delete
