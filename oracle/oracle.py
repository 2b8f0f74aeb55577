"""ctypes front end of the FP64 CPU oracle (oracle/rg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (recurgreen_amd)
never imports this module.

The C functions restate src/waveguide.f90, src/base_utils.f90 and
src/main.f90 of the reference; each wrapper names the file:line it follows.
Matrices are numpy complex128 row-major arrays; potentials keep the
reference's upot(ix, iy) layout (ix fastest), i.e. numpy shape (ny, nxdiso).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librgoracle.so")
_lib = None

DIRICHLET, PERIODIC = 0, 1
_BOUND = {"dirichlet": DIRICHLET, "periodic": PERIODIC}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, ip, dp, u32p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_uint32)
        L.rgo_create.restype = vp
        L.rgo_create.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 7 + [C.c_int, C.c_int, ip]
        L.rgo_destroy.argtypes = [vp]
        L.rgo_get_dims.argtypes = [vp, ip]
        L.rgo_get_modes.argtypes = [vp, dp, vp, vp, vp, vp, vp]
        L.rgo_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.rgo_normal_cell.restype = C.c_double
        L.rgo_normal_cell.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int]
        L.rgo_potential.argtypes = [vp, C.c_uint64, dp]
        L.rgo_project.restype = C.c_int
        L.rgo_project.argtypes = [vp, dp, C.c_int, vp, C.c_int]
        L.rgo_slice_matrix.argtypes = [vp, C.c_int, dp, vp, vp]
        L.rgo_set_projection.argtypes = [vp, C.c_int]
        L.rgo_inverse_alloc.restype = C.c_int
        L.rgo_inverse_alloc.argtypes = [C.c_int, vp]
        L.rgo_singular_values.restype = C.c_int
        L.rgo_singular_values.argtypes = [C.c_int, C.c_int, vp, dp]
        for name in ("rgo_transmission_matrix", "rgo_green_l0"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [vp, dp, vp]
        L.rgo_tdistrib_potential.restype = C.c_int
        L.rgo_tdistrib_potential.argtypes = [vp, dp, dp]
        L.rgo_tdistrib_seed.restype = C.c_int
        L.rgo_tdistrib_seed.argtypes = [vp, C.c_uint64, dp]
        L.rgo_tdistrib_batch.restype = C.c_int
        L.rgo_tdistrib_batch.argtypes = [vp, C.c_uint64, C.c_int, C.c_int, dp]
        for name in ("rgo_refl_trans_v1", "rgo_refl_trans_v2"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [vp, dp, vp, vp]
        L.rgo_green_column.restype = C.c_int
        L.rgo_green_column.argtypes = [vp, dp, C.c_int, vp]
        L.rgo_deposition.restype = C.c_int
        L.rgo_deposition.argtypes = [vp, dp, C.c_int, C.c_int, vp]
        L.rgo_solve_block_tridiag.restype = C.c_int
        L.rgo_solve_block_tridiag.argtypes = [C.c_int] * 3 + [vp] * 4
        L.rgo_bins.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dp]
        L.rgo_histogram.argtypes = [C.c_int64, dp, C.c_int, dp, dp, dp]
        L.rgo_normalize_histogram.argtypes = [C.c_int, dp, dp]
        L.rgo_moving_average.argtypes = [C.c_int, dp, C.c_int, dp]
        L.rgo_prepare_distrib.restype = C.c_double
        L.rgo_prepare_distrib.argtypes = [C.c_int64, dp, C.c_int, C.c_int, C.c_double, C.c_double, dp]
        L.rgo_omp_max_threads.restype = C.c_int
        L.rgo_wtime.restype = C.c_double
        if hasattr(L, "rgo_adjoint_grad"):
            L.rgo_adjoint_grad.restype = C.c_int
            L.rgo_adjoint_grad.argtypes = [vp, dp, C.c_int, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vp(a):
    return C.c_void_p(a.ctypes.data)


class OracleError(RuntimeError):
    pass


class Waveguide:
    """CPU restatement of waveguide_t (src/waveguide.f90:13-73, init :81-255).

    ny=None and dos_model=0 reproduce the reference (ny = nmu, mode-sum DOS);
    ny > nmu decouples the transverse grid (config mode), dos_model=1 uses the
    continuum DOS 1/(4 pi).
    """

    def __init__(self, boundtype="dirichlet", nxdiso=100, nxfree=0, wol=4.3, exfac=1.0, aratio=1.0,
                 dscat=0.2, dabso=0.0, napera=1.0, naperb=1.0, ny=None, dos_model=0):
        L = lib()
        st = C.c_int(0)
        self._h = L.rgo_create(_BOUND[boundtype], int(nxdiso), int(nxfree), float(wol), float(exfac),
                               float(aratio), float(dscat), float(dabso), float(napera), float(naperb),
                               int(ny or 0), int(dos_model), C.byref(st))
        if not self._h:
            raise OracleError(f"rgo_create failed with status {st.value}")
        d = (C.c_int * 11)()
        L.rgo_get_dims(self._h, d)
        (self.nmu, self.ny, self.nx, self.neigv, self.imin, self.imax, self.jmin, self.jmax,
         self.nprop, self.nxdiso, self.nxfree) = list(d)
        vals = np.zeros(4)
        n = self.nmu
        self.mu = np.zeros(n, complex)
        self.b = np.zeros(n, complex)
        self.abound = np.zeros(n, complex)
        self.abulk = np.zeros(n, complex)
        self.modal = np.zeros((self.ny, n), complex)
        L.rgo_get_modes(self._h, _dp(vals), _vp(self.mu), _vp(self.b), _vp(self.abound), _vp(self.abulk),
                        _vp(self.modal))
        self.kl, self.kdx, self.dos, self.stdu = vals.tolist()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.rgo_destroy(h)
            self._h = None

    def set_projection(self, method: int) -> None:
        """0: the reference's direct modal^H diag(u) modal product; 1: transform form (default)."""
        lib().rgo_set_projection(self._h, int(method))

    # -- disorder ---------------------------------------------------------------------------
    def potential(self, seed: int) -> np.ndarray:
        """upot(ix, iy) of realization `seed` as numpy (ny, nxdiso); src/waveguide.f90:366-372."""
        u = np.zeros((self.ny, self.nxdiso))
        lib().rgo_potential(self._h, int(seed), _dp(u))
        return u

    def project(self, u_slice: np.ndarray, method: int = 0) -> np.ndarray:
        u = np.ascontiguousarray(u_slice, dtype=np.float64)
        V = np.zeros((self.nmu, self.nmu), complex)
        st = lib().rgo_project(self._h, _dp(u), 1, _vp(V), int(method))
        if st:
            raise OracleError(f"rgo_project status {st}")
        return V

    def slice_matrix(self, ix: int, upot: np.ndarray | None) -> np.ndarray:
        A = np.zeros((self.nmu, self.nmu), complex)
        W = np.zeros_like(A)
        up = _dp(np.ascontiguousarray(upot)) if upot is not None else None
        keep = np.ascontiguousarray(upot) if upot is not None else None
        lib().rgo_slice_matrix(self._h, int(ix), _dp(keep) if keep is not None else up, _vp(A), _vp(W))
        return A

    # -- sweeps -----------------------------------------------------------------------------
    def _upot(self, upot):
        if upot is None:
            return None
        u = np.ascontiguousarray(upot, dtype=np.float64)
        assert u.shape == (self.ny, self.nxdiso), u.shape
        return u

    def transmission_matrix(self, upot) -> np.ndarray:
        u = self._upot(upot)
        t = np.zeros((self.imax - self.imin + 1, self.jmax - self.jmin + 1), complex)
        st = lib().rgo_transmission_matrix(self._h, _dp(u) if u is not None else None, _vp(t))
        if st:
            raise OracleError(f"sweep status {st}")
        return t

    def green_l0(self, upot) -> np.ndarray:
        u = self._upot(upot)
        G = np.zeros((self.nmu, self.nmu), complex)
        st = lib().rgo_green_l0(self._h, _dp(u) if u is not None else None, _vp(G))
        if st:
            raise OracleError(f"sweep status {st}")
        return G

    def tdistrib_potential(self, upot) -> np.ndarray:
        u = self._upot(upot)
        T = np.zeros(self.neigv)
        st = lib().rgo_tdistrib_potential(self._h, _dp(u) if u is not None else None, _dp(T))
        if st:
            raise OracleError(f"tdistrib status {st}")
        return T

    def tdistrib_seed(self, seed: int) -> np.ndarray:
        T = np.zeros(self.neigv)
        st = lib().rgo_tdistrib_seed(self._h, int(seed), _dp(T))
        if st:
            raise OracleError(f"tdistrib status {st}")
        return T

    def tdistrib_batch(self, seed0: int, nseed: int, nthreads: int = 1) -> np.ndarray:
        T = np.zeros(nseed * self.neigv)
        st = lib().rgo_tdistrib_batch(self._h, int(seed0), int(nseed), int(nthreads), _dp(T))
        if st:
            raise OracleError(f"tdistrib status {st}")
        return T

    def refl_trans(self, upot, version: int = 1):
        u = self._upot(upot)
        r = np.zeros((self.nmu, self.nmu), complex)
        t = np.zeros_like(r)
        fn = lib().rgo_refl_trans_v1 if version == 1 else lib().rgo_refl_trans_v2
        st = fn(self._h, _dp(u) if u is not None else None, _vp(r), _vp(t))
        if st:
            raise OracleError(f"refl_trans status {st}")
        return r, t

    def deposition(self, upot, fpos: int, ipos: int = 1) -> np.ndarray:
        """depmat (nmu, nmu) of compute_deposition (src/waveguide.f90:710-745), 1-based positions."""
        u = self._upot(upot)
        d = np.zeros((self.nmu, self.nmu), complex)
        st = lib().rgo_deposition(self._h, _dp(u) if u is not None else None, int(fpos), int(ipos), _vp(d))
        if st:
            raise OracleError(f"deposition status {st}")
        return d

    def green_column(self, upot, ipos: int = 1) -> np.ndarray:
        """G(x; ipos) for every slice x as (nx, nmu, nmu) (the block-tridiagonal solve)."""
        u = self._upot(upot)
        g = np.zeros((self.nx, self.nmu, self.nmu), complex)
        st = lib().rgo_green_column(self._h, _dp(u) if u is not None else None, int(ipos), _vp(g))
        if st:
            raise OracleError(f"green_column status {st}")
        return g

    def adjoint_grad(self, upot, loss: str = "g"):
        """FP64 adjoint: (loss value, d loss / d upot) with upot in U*dx/k units."""
        u = self._upot(upot)
        g = np.zeros_like(u)
        val = np.zeros(1)
        st = lib().rgo_adjoint_grad(self._h, _dp(u), {"g": 0, "tmax": 1}[loss], _dp(val), _dp(g))
        if st:
            raise OracleError(f"adjoint status {st}")
        return float(val[0]), g


# -- free functions ---------------------------------------------------------------------------
def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().rgo_philox4x32_10(c, k, o)
    return list(o)


def normal_cell(seed, ix, iy, stream=0):
    return lib().rgo_normal_cell(int(seed), int(stream), int(ix), int(iy))


def inverse(A: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(A, dtype=complex).copy()
    st = lib().rgo_inverse_alloc(X.shape[0], _vp(X))
    if st:
        raise OracleError(f"inverse status {st}")
    return X


def singular_values(A: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(A, dtype=complex)
    m, n = X.shape
    S = np.zeros(min(m, n))
    st = lib().rgo_singular_values(m, n, _vp(X), _dp(S))
    if st:
        raise OracleError(f"svd status {st}")
    return S


def solve_block_tridiag(amat, bdiag, imat):
    amat = np.ascontiguousarray(amat, dtype=complex)
    imat = np.ascontiguousarray(imat, dtype=complex)
    bdiag = np.ascontiguousarray(bdiag, dtype=complex)
    nb, bs, _ = amat.shape
    nrhs = imat.shape[2]
    x = np.zeros_like(imat)
    st = lib().rgo_solve_block_tridiag(nb, bs, nrhs, _vp(amat), _vp(bdiag), _vp(imat), _vp(x))
    if st:
        raise OracleError(f"block tridiag status {st}")
    return x


def bins(bintype: str, nbins: int, xmin=0.0, xmax=1.0):
    b = np.zeros(nbins + 1)
    lib().rgo_bins({"linear": 0, "chebyshev": 1}[bintype], nbins, float(xmin), float(xmax), _dp(b))
    return b


def histogram(samples, edges):
    s = np.ascontiguousarray(samples, dtype=np.float64)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    h = np.zeros(len(e) - 1)
    miss = C.c_double(0)
    lib().rgo_histogram(len(s), _dp(s), len(e) - 1, _dp(e), _dp(h), C.byref(miss))
    return h, miss.value


def moving_average(signal, n):
    s = np.ascontiguousarray(signal, dtype=np.float64)
    r = np.zeros(len(s) - n + 1)
    lib().rgo_moving_average(len(s), _dp(s), int(n), _dp(r))
    return r


def prepare_distrib(samples, nbins, bintype="chebyshev", xmin=0.0, xmax=1.0):
    s = np.ascontiguousarray(samples, dtype=np.float64)
    rho = np.zeros((nbins, 2))
    miss = lib().rgo_prepare_distrib(len(s), _dp(s), int(nbins), {"linear": 0, "chebyshev": 1}[bintype],
                                     float(xmin), float(xmax), _dp(rho))
    return rho, miss


def omp_max_threads():
    return lib().rgo_omp_max_threads()


def wtime():
    return lib().rgo_wtime()
