"""ctypes binding of the C ABI declared in include/recurgreen_hip.h.

The shared library is built in-tree (recurgreen_amd/librecurgreen_hip.so, see build.py).
There is no fallback: if the library is missing or cannot be loaded, every entry point
raises RecurGreenError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librecurgreen_hip.so")

RG_OK, RG_EINVAL, RG_ESINGULAR, RG_ENOCONV, RG_ENOMEM, RG_EDEVICE, RG_ENOTSUP, RG_ENONFINITE = range(8)
PRECISIONS = {"fp64": 0, "mixed_fp32": 1, "mixed_bf16x3": 2}
BOUNDTYPES = {"dirichlet": 0, "periodic": 1}

# Every symbol include/recurgreen_hip.h declares.
EXPORTS = (
    "rg_plan_create", "rg_plan_destroy", "rg_plan_dims", "rg_tdistrib", "rg_tdistrib_device",
    "rg_tdistrib_potential", "rg_transmission_potential", "rg_refl_trans", "rg_refl_trans_potential", "rg_deposition", "rg_deposition_potential", "rg_tprofile_reserve", "rg_tprofile", "rg_tprofile_potential", "rg_wavefunction", "rg_wavefunction_potential",
    "rg_tdistrib_uneven",
    "rg_adjoint_reserve", "rg_adjoint_potential", "rg_adjoint_potential_device",
    "rg_generate_potential", "rg_histogram", "rg_histogram_noalloc",
    "rg_plan_stream", "rg_plan_synchronize", "rg_plan_profile", "rg_plan_kernel_times",
    "rg_plan_set_option", "rg_plan_get_option", "rg_status_string", "rg_version",
)


class RecurGreenError(RuntimeError):
    """Raised for a non-zero rg_status; .code holds the status."""

    def __init__(self, code: int, msg: str | None = None):
        self.code = int(code)
        super().__init__(msg or status_string(code))


class rg_params(C.Structure):
    _fields_ = [
        ("boundtype", C.c_int32), ("nxdiso", C.c_int32), ("nxfree", C.c_int32),
        ("wol", C.c_double), ("exfac", C.c_double), ("aratio", C.c_double), ("dscat", C.c_double),
        ("dabso", C.c_double), ("napera", C.c_double), ("naperb", C.c_double),
        ("ny", C.c_int32), ("dos_model", C.c_int32),
    ]


_lib = None


def lib():
    """Load librecurgreen_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RecurGreenError(RG_EDEVICE, f"{LIB_PATH} is not built; run recurgreen_amd.build.build()")
    # One HIP runtime per process: when PyTorch-ROCm is installed, let it load and initialise its
    # runtime first, so that this library binds to the same libamdhip64 and device pointers pass
    # freely between torch tensors and the plan (a second runtime loaded later hides the GPUs
    # from torch).
    try:
        import torch

        torch.cuda.is_available()
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    i32p, i64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    sig = {
        "rg_plan_create": [C.POINTER(rg_params), i32, i32, C.POINTER(vp)],
        "rg_plan_destroy": [vp],
        "rg_plan_dims": [vp, i32p, dp],
        "rg_tdistrib": [vp, i64, i32, i32, vp, vp],
        "rg_tdistrib_device": [vp, i64, i32, i32, vp, vp],
        "rg_tdistrib_potential": [vp, vp, i32, i32, vp, vp],
        "rg_transmission_potential": [vp, vp, i32, i32, vp, vp],
        "rg_refl_trans": [vp, i64, i32, vp, vp, vp],
        "rg_adjoint_reserve": [vp, i32],
        "rg_adjoint_potential": [vp, vp, i32, i32, vp, vp, vp],
        "rg_adjoint_potential_device": [vp, vp, i32, i32, vp, vp, vp],
        "rg_refl_trans_potential": [vp, vp, i32, vp, vp, vp],
        "rg_deposition": [vp, i64, i32, i32, i32, vp, vp],
        "rg_deposition_potential": [vp, vp, i32, i32, i32, vp, vp],
        "rg_tprofile_reserve": [vp, i32],
        "rg_tprofile": [vp, i64, i32, i32, vp, vp, vp],
        "rg_tprofile_potential": [vp, vp, i32, i32, vp, vp, vp],
        "rg_wavefunction": [vp, i64, i32, i32, vp, vp],
        "rg_wavefunction_potential": [vp, vp, i32, i32, vp, vp],
        "rg_generate_potential": [vp, i64, i32, vp],
        "rg_histogram": [vp, vp, i64, i32, dp, i32, i64p, i64p, i64p],
        "rg_histogram_noalloc": [i32, vp, vp, i64, dp, i32, i64p, i64p, i64p],
        "rg_plan_stream": [vp, C.POINTER(vp)],
        "rg_plan_synchronize": [vp],
        "rg_plan_profile": [vp, i32],
        "rg_plan_kernel_times": [vp, i32, C.c_char_p, dp, i64p, i32p],
        "rg_plan_set_option": [vp, C.c_char_p, i32],
        "rg_tdistrib_uneven": [C.POINTER(rg_params), C.POINTER(C.c_int64), i32, i32, i32, i32, vp, i64p, i32p],
        "rg_plan_get_option": [vp, C.c_char_p, i32p],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.rg_status_string.argtypes = [C.c_int32]
    L.rg_status_string.restype = C.c_char_p
    L.rg_version.argtypes = []
    L.rg_version.restype = C.c_char_p
    _lib = L
    return L


def status_string(code: int) -> str:
    try:
        return lib().rg_status_string(int(code)).decode()
    except RecurGreenError:
        return f"RecurGreen status {code}"


def check(code: int) -> None:
    if code != RG_OK:
        raise RecurGreenError(code)


def ptr(a) -> C.c_void_p:
    """Address of a numpy array or a torch tensor (device or host)."""
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, np.ndarray):
        return C.c_void_p(a.ctypes.data)
    return C.c_void_p(int(a.data_ptr()))


def make_params(params: dict) -> rg_params:
    """rg_params from the keys of the &waveguide namelist (plus ny, dos_model)."""
    return rg_params(
        boundtype=BOUNDTYPES[params.get("boundtype", "dirichlet")],
        nxdiso=int(params["nxdiso"]), nxfree=int(params.get("nxfree", 0)),
        wol=float(params["wol"]), exfac=float(params.get("exfac", 1.0)),
        aratio=float(params.get("aratio", 1.0)), dscat=float(params.get("dscat", 0.0)),
        dabso=float(params.get("dabso", 0.0)), napera=float(params.get("napera", 1.0)),
        naperb=float(params.get("naperb", 1.0)), ny=int(params.get("ny") or 0),
        dos_model=int(params.get("dos_model", 0)),
    )


def tdistrib_uneven(params: list, seeds, precision: str = "fp64", device: int = 0, batch_max: int = 4096):
    """rg_tdistrib_uneven: one call over realizations with their own parameters and seeds.
    Returns (list of per-realization eigenvalue arrays in the given order, status array, code)."""
    n = len(params)
    arr = (rg_params * max(1, n))(*[make_params(p) for p in params])
    sd = np.ascontiguousarray(seeds, dtype=np.int64)
    if sd.shape != (n,):
        raise ValueError("one seed per realization")
    off = np.zeros(n + 1, np.int64)
    st = np.zeros(max(1, n), np.int32)
    # offsets first (eigenvalue counts), then the sweep: the counts follow from init alone
    total = 0
    for p in params:
        probe = make_params(p)
        total += _neigv_of(probe)
    out = np.zeros(max(1, total), np.float64)
    code = lib().rg_tdistrib_uneven(arr, sd.ctypes.data_as(C.POINTER(C.c_int64)), n, PRECISIONS[precision],
                                    int(device), int(batch_max), out.ctypes.data, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                    st.ctypes.data_as(C.POINTER(C.c_int32)))
    vals = [out[off[r]:off[r + 1]].copy() for r in range(n)]
    return vals, st[:n], code


def histogram_noalloc(samples, bins: np.ndarray, device: int = 0, stream: int = 0):
    """rg_histogram_noalloc: (counts, missing, dropped) of device-resident samples (a CUDA tensor of
    float64 on `device`) on the edges `bins`, with no plan and no allocation in the library; `stream`
    is a hipStream_t handle (0: the null stream)."""
    b = np.ascontiguousarray(bins, dtype=np.float64)
    nb = len(b) - 1
    counts = np.zeros(max(nb, 1), np.int64)
    miss, drop = C.c_int64(0), C.c_int64(0)
    check(lib().rg_histogram_noalloc(int(device), C.c_void_p(int(stream)), ptr(samples), int(samples.numel()),
                                     b.ctypes.data_as(C.POINTER(C.c_double)), nb,
                                     counts.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(miss), C.byref(drop)))
    return counts[:nb], miss.value, drop.value


def _neigv_of(p: rg_params) -> int:
    """Eigenvalue count of one waveguide (src/waveguide.f90:208 Dirichlet, :165 periodic), or 0
    when the parameters are rejected (the library reports the status)."""
    import math

    m = min(p.napera, p.naperb)
    if p.boundtype == BOUNDTYPES["dirichlet"]:
        return max(0, int(math.floor(2.0 * m * p.wol)))
    return max(0, 2 * int(math.floor(m * p.wol)) + 1)


class Plan:
    """Owner of one rg_plan (device buffers, stream, mode constants)."""

    def __init__(self, params: dict, batch_max: int = 256, device: int = 0):
        L = lib()
        p = make_params(params)
        h = C.c_void_p()
        check(L.rg_plan_create(C.byref(p), int(batch_max), int(device), C.byref(h)))
        self._h = h
        self.batch_max = int(batch_max)
        self.device = int(device)
        d = (C.c_int32 * 11)()
        v = (C.c_double * 4)()
        check(L.rg_plan_dims(self._h, d, v))
        (self.nmu, self.ny, self.nx, self.neigv, self.imin, self.imax, self.jmin, self.jmax,
         self.nprop, self.nxdiso, self.nxfree) = list(d)
        self.kl, self.kdx, self.dos, self.stdu = list(v)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().rg_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _upot(self, upot: np.ndarray) -> tuple:
        """upot(ix, iy[, r]) as a Fortran-ordered float64 array and its realization count; raises
        RecurGreenError(RG_EINVAL) unless the leading dimensions are (nxdiso, ny), so that the
        library never reads past the caller's buffer."""
        u = np.asfortranarray(upot, dtype=np.float64)
        if u.ndim not in (2, 3) or u.shape[:2] != (self.nxdiso, self.ny) or (u.ndim == 3 and u.shape[2] < 1):
            raise RecurGreenError(RG_EINVAL, f"potential of shape {np.shape(upot)}: expected (nxdiso, ny[, nreal]) = "
                                             f"({self.nxdiso}, {self.ny}[, nreal])")
        return u, (1 if u.ndim == 2 else u.shape[2])

    # -- sweeps ---------------------------------------------------------------------------
    def tdistrib(self, seed0: int, nseed: int, precision: str = "fp64"):
        T = np.zeros(nseed * self.neigv)
        st = np.zeros(nseed, np.int32)
        code = lib().rg_tdistrib(self._h, int(seed0), int(nseed), PRECISIONS[precision], ptr(T), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENOCONV, RG_ENONFINITE):
            check(code)
        return T, st, code

    def tdistrib_device(self, seed0: int, nseed: int, precision: str, out, status=None):
        check(lib().rg_tdistrib_device(self._h, int(seed0), int(nseed), PRECISIONS[precision], ptr(out),
                                       ptr(status)))

    def tdistrib_potential(self, upot: np.ndarray, precision: str = "fp64"):
        u, nreal = self._upot(upot)
        T = np.zeros(nreal * self.neigv)
        st = np.zeros(nreal, np.int32)
        code = lib().rg_tdistrib_potential(self._h, ptr(u), nreal, PRECISIONS[precision], ptr(T), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENOCONV, RG_ENONFINITE):
            check(code)
        return T, st, code

    def transmission_potential(self, upot: np.ndarray | None, nreal: int = 1, precision: str = "fp64"):
        m, n = self.imax - self.imin + 1, self.jmax - self.jmin + 1
        if upot is not None:
            u, nreal = self._upot(upot)
        else:
            u = None
        t = np.zeros((nreal, n, m), complex)  # column-major m x n per realization
        st = np.zeros(nreal, np.int32)
        code = lib().rg_transmission_potential(self._h, ptr(u), int(nreal), PRECISIONS[precision], ptr(t), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENONFINITE):
            check(code)
        return np.ascontiguousarray(t.transpose(0, 2, 1)), st, code

    def refl_trans(self, upot: np.ndarray | None = None, seed0: int = 1, nseed: int = 1):
        """(r, t, status, code) of compute_refl_trans_v1 (src/waveguide.f90:469-519), complex128:
        r and t are (nreal, nmu, nmu); explicit potentials upot(ix, iy[, r]) or seeds seed0...."""
        n = self.nmu
        if upot is not None:
            u, nreal = self._upot(upot)
        else:
            u, nreal = None, int(nseed)
        r = np.zeros((nreal, n, n), complex)  # column-major per realization
        t = np.zeros((nreal, n, n), complex)
        st = np.zeros(nreal, np.int32)
        if u is None:
            code = lib().rg_refl_trans(self._h, int(seed0), nreal, ptr(r), ptr(t), ptr(st))
        else:
            code = lib().rg_refl_trans_potential(self._h, ptr(u), nreal, ptr(r), ptr(t), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENONFINITE):
            check(code)
        return (np.ascontiguousarray(r.transpose(0, 2, 1)), np.ascontiguousarray(t.transpose(0, 2, 1)), st, code)

    def deposition(self, fpos: int, ipos: int = 1, upot: np.ndarray | None = None, seed0: int = 1, nseed: int = 1):
        """(dep, status, code) of compute_deposition (src/waveguide.f90:710-745), complex128: dep is
        (nreal, nmu, nmu), 1-based slice positions; explicit potentials upot(ix, iy[, r]) or seeds."""
        n = self.nmu
        if upot is not None:
            u, nreal = self._upot(upot)
        else:
            u, nreal = None, int(nseed)
        d = np.zeros((nreal, n, n), complex)  # column-major per realization
        st = np.zeros(nreal, np.int32)
        if u is None:
            code = lib().rg_deposition(self._h, int(seed0), nreal, int(fpos), int(ipos), ptr(d), ptr(st))
        else:
            code = lib().rg_deposition_potential(self._h, ptr(u), nreal, int(fpos), int(ipos), ptr(d), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENONFINITE):
            check(code)
        return np.ascontiguousarray(d.transpose(0, 2, 1)), st, code

    def tprofile_reserve(self, nreal_max: int) -> None:
        check(lib().rg_tprofile_reserve(self._h, int(nreal_max)))

    def tprofile_pass_size(self, nreal: int, fraction: float = 0.6) -> int:
        """Realizations per tprofile pass: at most nreal and batch_max, and as many as fit in `fraction`
        of the device's free memory (every slice's Gss and Gs1 plus five work matrices, complex128,
        and the intensities per realization)."""
        np_ = (self.nmu + 15) // 16 * 16
        per = (2 * self.nx + 5) * np_ * np_ * 16 + self.nx * self.nmu * 8
        cap = max(1, min(int(nreal), self.batch_max))
        try:
            import torch

            free = torch.cuda.mem_get_info(self.device)[0]
            cap = max(1, min(cap, int(fraction * free) // per))
        except Exception:
            pass
        return cap

    def tprofile(self, upot: np.ndarray | None = None, seed0: int = 1, nseed: int = 1, filtered: bool = False,
                 reserve: int | None = None):
        """(tvals (nreal, nv) descending, intensity (nreal, nx, nv), status, code): the eigenvalues of
        t^H t and |G(x;1) s U(:,v)|^2 per slice and eigenvector (compute_tprofile,
        src/waveguide.f90:843-907, before the window sum; filtered: t over the apertures as
        compute_tprofile_group, :951-954).  Reserves the workspace for min(nreal, reserve or
        batch_max) realizations per pass."""
        if upot is not None:
            u, nreal = self._upot(upot)
        else:
            u, nreal = None, int(nseed)
        self.tprofile_reserve(max(1, reserve or self.tprofile_pass_size(nreal)))
        nv = (self.jmax - self.jmin + 1) if filtered else self.nmu
        tv = np.zeros((nreal, nv))
        it = np.zeros((nreal, self.nx, nv))
        st = np.zeros(nreal, np.int32)
        f = int(bool(filtered))
        if u is None:
            code = lib().rg_tprofile(self._h, int(seed0), nreal, f, ptr(tv), ptr(it), ptr(st))
        else:
            code = lib().rg_tprofile_potential(self._h, ptr(u), nreal, f, ptr(tv), ptr(it), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENOCONV, RG_ENONFINITE):
            check(code)
        return tv, it, st, code

    def wavefunction(self, ipos: int, imode: int, upot: np.ndarray | None = None, seed: int = 1):
        """(psi (nx, ny) complex128, status, code): compute_wavefunction's field
        (src/waveguide.f90:636-697) for one realization, explicit upot(ix, iy) or a seed."""
        self.tprofile_reserve(1)
        psi = np.zeros((self.nx, self.ny), complex)
        st = np.zeros(1, np.int32)
        if upot is None:
            code = lib().rg_wavefunction(self._h, int(seed), int(ipos), int(imode), ptr(psi), ptr(st))
        else:
            u, nreal = self._upot(upot)
            if nreal != 1:
                raise RecurGreenError(RG_EINVAL, "one potential (nxdiso, ny) per wavefunction")
            code = lib().rg_wavefunction_potential(self._h, ptr(u), int(ipos), int(imode), ptr(psi), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENONFINITE):
            check(code)
        return psi, st, code

    LOSSES = {"g": 0, "tmax": 1}

    def adjoint_reserve(self, nreal_max: int) -> None:
        check(lib().rg_adjoint_reserve(self._h, int(nreal_max)))

    def adjoint(self, upot: np.ndarray, loss: str = "g"):
        """(loss[r], dL/du as (nxdiso, ny, nreal) Fortran-ordered, status, code) for explicit host
        potentials; complex128 adjoint sweep."""
        u, nreal = self._upot(upot)
        if u.ndim == 2:
            u = u[:, :, None]
        self.adjoint_reserve(min(nreal, self.batch_max))
        val = np.zeros(nreal)
        grad = np.zeros((nreal, self.ny, self.nxdiso))
        st = np.zeros(nreal, np.int32)
        code = lib().rg_adjoint_potential(self._h, ptr(u), nreal, self.LOSSES[loss], ptr(val), ptr(grad), ptr(st))
        if code not in (RG_OK, RG_ESINGULAR, RG_ENONFINITE):
            check(code)
        return val, grad.transpose(2, 1, 0), st, code

    def adjoint_device(self, upot, loss: str = "g"):
        """Device variant on torch CUDA tensors: upot is (nreal, ny, nxdiso) float64 contiguous (the
        (nxdiso, ny, nreal) Fortran layout).  Returns (loss, grad) tensors; asynchronous on stream()."""
        import torch

        if (not isinstance(upot, torch.Tensor) or upot.dtype != torch.float64 or not upot.is_cuda or upot.dim() != 3
                or tuple(upot.shape[1:]) != (self.ny, self.nxdiso) or upot.shape[0] < 1 or not upot.is_contiguous()):
            raise RecurGreenError(RG_EINVAL, f"adjoint_device expects a contiguous float64 CUDA tensor of shape "
                                             f"(nreal, {self.ny}, {self.nxdiso})")
        nreal = upot.shape[0]
        self.adjoint_reserve(min(nreal, self.batch_max))
        val = torch.empty(nreal, dtype=torch.float64, device=upot.device)
        grad = torch.empty_like(upot)
        check(lib().rg_adjoint_potential_device(self._h, ptr(upot), nreal, self.LOSSES[loss], ptr(val), ptr(grad),
                                                None))
        return val, grad

    def generate_potential(self, seed0: int, nreal: int) -> np.ndarray:
        """upot(ix, iy, r) as a Fortran-ordered (nxdiso, ny, nreal) array."""
        u = np.zeros((nreal, self.ny, self.nxdiso))
        check(lib().rg_generate_potential(self._h, int(seed0), int(nreal), ptr(u)))
        return u.transpose(2, 1, 0)

    def histogram(self, samples, bins: np.ndarray, on_device: bool | None = None):
        b = np.ascontiguousarray(bins, dtype=np.float64)
        nb = len(b) - 1
        counts = np.zeros(nb, np.int64)
        miss, drop = C.c_int64(0), C.c_int64(0)
        if on_device is None:
            on_device = not isinstance(samples, np.ndarray)
        if isinstance(samples, np.ndarray):
            s = np.ascontiguousarray(samples, dtype=np.float64)
            n = s.size
        else:
            s = samples
            n = samples.numel()
        check(lib().rg_histogram(self._h, ptr(s), int(n), int(bool(on_device)),
                                 b.ctypes.data_as(C.POINTER(C.c_double)), nb,
                                 counts.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(miss), C.byref(drop)))
        return counts, miss.value, drop.value

    # -- stream / timing ------------------------------------------------------------------
    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().rg_plan_stream(self._h, C.byref(s)))
        return int(s.value or 0)

    def synchronize(self):
        check(lib().rg_plan_synchronize(self._h))

    # -- launch options (the autotuner's choices) -------------------------------------------
    OPTIONS = ("inverse_fp64", "inverse_mixed", "dyson", "dyson_fp64", "project_parts", "project_wide_min", "refine", "refine_cap", "lanes",
               "eigen", "graph", "adjoint_precision")

    def set_option(self, key: str, value: int) -> None:
        check(lib().rg_plan_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key: str) -> int:
        v = C.c_int32(0)
        check(lib().rg_plan_get_option(self._h, key.encode(), C.byref(v)))
        return v.value

    def options(self) -> dict:
        return {k: self.get_option(k) for k in self.OPTIONS}

    def profile(self, enable: bool):
        check(lib().rg_plan_profile(self._h, int(bool(enable))))

    def kernel_times(self) -> dict:
        cap = 16
        names = C.create_string_buffer(32 * cap)
        ms = (C.c_double * cap)()
        nl = (C.c_int64 * cap)()
        cnt = C.c_int32(0)
        check(lib().rg_plan_kernel_times(self._h, cap, names, ms, nl, C.byref(cnt)))
        out = {}
        raw = names.raw
        for k in range(cnt.value):
            nm = raw[32 * k:32 * k + 32].split(b"\0", 1)[0].decode()
            out[nm] = {"ms": ms[k], "launches": nl[k]}
        return out
