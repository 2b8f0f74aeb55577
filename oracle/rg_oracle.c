/*
 * rg_oracle.c -- FP64 CPU oracle for the recursive Green function (RGF) sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against; it is never part of the product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * It restates, in plain C99 double-complex arithmetic, the algorithm of the
 * reference's Fortran/LAPACK path (RecurGreen v1.0):
 *   - mode set-up                 src/waveguide.f90:81-255   (rgo_create)
 *   - slice matrices + disorder   src/waveguide.f90:349-384  (rgo_slice_matrix)
 *   - RGF sweep, Fisher-Lee, SVD  src/waveguide.f90:392-462  (rgo_tdistrib_potential)
 *   - r/t sweep, version 1        src/waveguide.f90:469-519  (rgo_refl_trans_v1)
 *   - r/t by block-tridiag solve  src/waveguide.f90:526-553,
 *                                 src/base_utils.f90:296-335 (rgo_refl_trans_v2)
 *   - deposition matrix           src/waveguide.f90:710-745  (rgo_deposition)
 *   - Green function column block src/waveguide.f90:873-878  (rgo_green_column)
 *   - zgetrf+zgetri inverse       src/base_utils.f90:235-274 (rgo_inverse)
 *   - zgesvd singular values      src/base_utils.f90:341-380 (rgo_singular_values,
 *                                 one-sided Jacobi instead of LAPACK)
 *   - histogram chain             src/base_utils.f90:468-570, src/main.f90:223-251
 *   - OpenMP realization loop     src/main.f90:134-179       (rgo_tdistrib_batch)
 *
 * The disorder generator is NOT the reference's (gfortran's random_number
 * stream, src/random_normal.f90:40-67, is compiler-runtime specific); both the
 * oracle and the GPU library use the counter-based Philox4x32-10 generator of
 * Salmon et al. (SC'11) keyed by (seed, ix, iy / 2), followed by the reference's
 * Box-Muller transform (src/random_normal.f90:63-65), both of its outputs: the
 * cosine one for even iy, the sine one for odd iy.
 *
 * Conventions: matrices are row-major, element (i, j) of an m x n matrix at
 * [i*n + j]; indices inside the code are 0-based, comments cite the
 * reference's 1-based formulas.  The potential array upot(ix, iy) keeps the
 * reference layout (src/waveguide.f90:43): ix fastest, upot[iy*nxdiso + ix],
 * in units of U*dx/k.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "rg_oracle.h"

/* ---------------------------------------------------------------------------------------------
 * Set-up.  src/waveguide.f90:81-255.  Extensions over the reference (both documented):
 *   ny_override > 0 decouples the transverse grid from the mode count (the reference forces
 *   ny = nmu at :162/:202); dos_model = 1 uses the continuum DOS 1/(4 pi) instead of the
 *   finite mode sum of :243.  ny_override = 0 and dos_model = 0 reproduce the reference.
 * ------------------------------------------------------------------------------------------- */
void rgo_destroy(rgo_wg *w);

rgo_wg *rgo_create(int boundtype, int nxdiso, int nxfree, double wol, double exfac, double aratio,
                   double dscat, double dabso, double napera, double naperb, int ny_override,
                   int dos_model, int *status)
{
    *status = RGO_EINVAL;
    if (!(wol >= 2.220446049250313e-16) || !(aratio >= 2.220446049250313e-16)) return NULL; /* :91-96 */
    if (exfac < 1.0 || exfac > 1e4) return NULL;                                             /* :97 */
    if (nxfree < 0 || nxdiso < 1) return NULL;                                                 /* :112 */
    if (dscat < 0.0 || dabso < 0.0) return NULL;                                               /* :121 */
    if (napera <= 0.0 || napera > 1.0 || naperb <= 0.0 || naperb > 1.0) return NULL;           /* :132-140 */
    if (nxdiso + 2 * nxfree < 2) return NULL; /* the reference writes amat(2,:,:) (:360) */

    rgo_wg *w = (rgo_wg *)calloc(1, sizeof(rgo_wg));
    if (!w) { *status = RGO_ENOMEM; return NULL; }
    w->boundtype = boundtype;
    w->nxdiso = nxdiso; w->nxfree = nxfree; w->nx = nxdiso + 2 * nxfree;  /* :117 */
    w->wol = wol; w->exfac = exfac; w->aratio = aratio;
    w->dscat = dscat; w->dabso = dabso; w->napera = napera; w->naperb = naperb;
    w->kl = 2.0 * RGO_PI * wol * aratio;                                  /* :104 */
    w->kdx = w->kl / nxdiso;                                               /* :118 */

    if (boundtype == RGO_DIRICHLET) {
        if (fmod(wol, 0.5) == 0.0) { free(w); return NULL; }              /* :194 */
        w->nmu = (int)floor(2.0 * exfac * wol);                            /* :201 */
        w->imin = 1; w->imax = (int)floor(2.0 * naperb * wol);            /* :203-206 */
        w->jmin = 1; w->jmax = (int)floor(2.0 * napera * wol);
        w->neigv = (int)floor(2.0 * fmin(napera, naperb) * wol);           /* :208 */
        w->nprop = (int)floor(2.0 * wol);                                  /* :209 */
    } else if (boundtype == RGO_PERIODIC) {
        if (fmod(wol, 1.0) == 0.0) { free(w); return NULL; }              /* :154 */
        w->nmu = 2 * (int)floor(exfac * wol) + 1;                          /* :161 */
        w->imin = w->nmu / 2 + 1 - (int)floor(naperb * wol);               /* :163-166 */
        w->imax = w->nmu / 2 + 1 + (int)floor(naperb * wol);
        w->jmin = w->nmu / 2 + 1 - (int)floor(napera * wol);
        w->jmax = w->nmu / 2 + 1 + (int)floor(napera * wol);
        w->neigv = 2 * (int)floor(fmin(napera, naperb) * wol) + 1;         /* :168 */
        w->nprop = 2 * (int)floor(wol) + 1;                                /* :169 */
    } else { free(w); return NULL; }
    w->ny = ny_override > 0 ? ny_override : w->nmu;                        /* :162/:202 */
    if (w->ny < w->nmu || w->nmu < 1) { free(w); return NULL; }

    int n = w->nmu, ny = w->ny;
    w->kpmesh = (double *)malloc(sizeof(double) * n);
    w->ymesh = (double *)malloc(sizeof(double) * ny);
    w->mumesh = (zc *)malloc(sizeof(zc) * n);
    w->bmat = (zc *)malloc(sizeof(zc) * n);
    w->abound = (zc *)malloc(sizeof(zc) * n);
    w->abulk = (zc *)malloc(sizeof(zc) * n);
    w->modal = (zc *)malloc(sizeof(zc) * (size_t)ny * n);
    if (!w->kpmesh || !w->ymesh || !w->mumesh || !w->bmat || !w->abound || !w->abulk || !w->modal) {
        rgo_destroy(w); *status = RGO_ENOMEM; return NULL;
    }
    for (int i = 0; i < ny; i++) w->ymesh[i] = (i + 1 - 0.5) / ny;         /* :173/:213 */
    for (int m = 0; m < n; m++) {
        if (boundtype == RGO_DIRICHLET) w->kpmesh[m] = 0.5 * (m + 1) / wol; /* :218 */
        else w->kpmesh[m] = (double)(m + 1 - n / 2 - 1) / wol;               /* :178 */
    }
    if (boundtype == RGO_DIRICHLET) {
        double nrm = sqrt(2.0 / ny);                                        /* :228 */
        for (int i = 0; i < ny; i++)
            for (int m = 0; m < n; m++)
                w->modal[(size_t)i * n + m] = nrm * sin(2.0 * RGO_PI * wol * w->kpmesh[m] * w->ymesh[i]); /* :225 */
    } else {
        double nrm = 1.0 / sqrt((double)ny);                                /* :188 */
        for (int i = 0; i < ny; i++)
            for (int m = 0; m < n; m++)
                w->modal[(size_t)i * n + m] = nrm * cexp(I * 2.0 * RGO_PI * wol * w->kpmesh[m] * w->ymesh[i]); /* :185 */
    }
    double dsum = 0.0;
    for (int m = 0; m < n; m++) {
        double kp2 = w->kpmesh[m] * w->kpmesh[m];
        w->mumesh[m] = csqrt(1.0 + I * dabso / w->kl - kp2);               /* :241 */
        dsum += creal(1.0 / csqrt(CMPLX(1.0, 0.0) - kp2));                 /* :243 */
    }
    w->dos = dos_model == 1 ? 1.0 / (4.0 * RGO_PI) : dsum / (wol * (2.0 * RGO_PI) * (2.0 * RGO_PI));
    for (int m = 0; m < n; m++) {
        zc mu = w->mumesh[m];
        zc th = mu * w->kdx;
        w->bmat[m] = mu / csin(th);                                         /* :248 */
        w->abound[m] = I * mu - mu / ctan(th);                              /* :358 */
        w->abulk[m] = -2.0 * mu / ctan(th);                                 /* :360 */
    }
    w->stdu = sqrt(dscat * ny / (2.0 * RGO_PI * RGO_PI * w->dos * wol * nxdiso)); /* :370 */
    w->proj_method = 1;
    *status = RGO_OK;
    return w;
}

void rgo_destroy(rgo_wg *w)
{
    if (!w) return;
    free(w->kpmesh); free(w->ymesh); free(w->mumesh); free(w->bmat);
    free(w->abound); free(w->abulk); free(w->modal); free(w);
}

/* dims[] = {nmu, ny, nx, neigv, imin, imax, jmin, jmax, nprop, nxdiso, nxfree} */
void rgo_get_dims(const rgo_wg *w, int *dims)
{
    dims[0] = w->nmu; dims[1] = w->ny; dims[2] = w->nx; dims[3] = w->neigv;
    dims[4] = w->imin; dims[5] = w->imax; dims[6] = w->jmin; dims[7] = w->jmax;
    dims[8] = w->nprop; dims[9] = w->nxdiso; dims[10] = w->nxfree;
}

/* vals[] = {kl, kdx, dos, stdu} ; mu, b, abound, abulk: nmu complex each (may be NULL) */
void rgo_get_modes(const rgo_wg *w, double *vals, zc *mu, zc *b, zc *abound, zc *abulk, zc *modal)
{
    if (vals) { vals[0] = w->kl; vals[1] = w->kdx; vals[2] = w->dos; vals[3] = w->stdu; }
    size_t n = (size_t)w->nmu;
    if (mu) memcpy(mu, w->mumesh, sizeof(zc) * n);
    if (b) memcpy(b, w->bmat, sizeof(zc) * n);
    if (abound) memcpy(abound, w->abound, sizeof(zc) * n);
    if (abulk) memcpy(abulk, w->abulk, sizeof(zc) * n);
    if (modal) memcpy(modal, w->modal, sizeof(zc) * n * (size_t)w->ny);
}

/* ---------------------------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): 10 rounds of the 4x32 Philox S-box with Weyl key schedule.
 * ------------------------------------------------------------------------------------------- */
void rgo_philox4x32_10(const uint32_t *ctr_in, const uint32_t *key_in, uint32_t *out)
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Key and counter of the disorder cells (ix, 2p) and (ix, 2p + 1) of realization `seed` (0-based ix, p).
 * stream 0 = disorder potential. */
static void rgo_cell_key(uint64_t seed, uint32_t stream, int ix, int p, uint32_t *ctr, uint32_t *key)
{
    ctr[0] = (uint32_t)p; ctr[1] = (uint32_t)ix; ctr[2] = stream; ctr[3] = 0u;
    key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32) ^ 0x52474E31u;
}

/* Two 53-bit uniforms from one Philox block, then the Box-Muller transform (the reference keeps its cosine
 * output only, src/random_normal.f90:64; both outputs are used here, for two cells):
 * sqrt(-2 log u1) cos(2 pi u2) (odd = 0) or sqrt(-2 log u1) sin(2 pi u2) (odd = 1), u1 in (0,1], u2 in [0,1). */
double rgo_normal_from_bits(const uint32_t *r, int odd)
{
    uint64_t a = ((uint64_t)(r[0] >> 5) << 26) | (uint64_t)(r[1] >> 6);
    uint64_t b = ((uint64_t)(r[2] >> 5) << 26) | (uint64_t)(r[3] >> 6);
    double u1 = (double)(a + 1) * 0x1.0p-53;
    double u2 = (double)b * 0x1.0p-53;
    double th = 2.0 * RGO_PI * u2;
    return sqrt(-2.0 * log(u1)) * (odd ? sin(th) : cos(th));
}

/* Cell (ix, iy): the Philox block of the pair p = iy / 2, the cosine output for even iy, the sine for odd. */
double rgo_normal_cell(uint64_t seed, uint32_t stream, int ix, int iy)
{
    uint32_t ctr[4], key[2], r[4];
    rgo_cell_key(seed, stream, ix, iy >> 1, ctr, key);
    rgo_philox4x32_10(ctr, key, r);
    return rgo_normal_from_bits(r, iy & 1);
}

/* upot(ix, iy) = stdu * xi(seed, ix, iy): src/waveguide.f90:366-372 with the Philox stream. */
void rgo_potential(const rgo_wg *w, uint64_t seed, double *upot)
{
    for (int iy = 0; iy < w->ny; iy++)
        for (int ix = 0; ix < w->nxdiso; ix++)
            upot[(size_t)iy * w->nxdiso + ix] = w->stdu * rgo_normal_cell(seed, 0u, ix, iy);
}

/* ---------------------------------------------------------------------------------------------
 * Projection of one potential slice u(iy) onto the modes: V = modal^H diag(u) modal
 * (src/waveguide.f90:377-380).  method 0: the reference's direct product, O(N^2 ny).
 * method 1, O(N ny): Dirichlet, chi_m(y) = sqrt(2/ny) sin(m pi y) and
 * sin(a) sin(b) = [cos(a-b) - cos(a+b)]/2 give V_mn = C(|m-n|) - C(m+n) with
 * C(q) = (1/ny) sum_i u_i cos(q pi y_i), q = 0..2N (Toeplitz minus Hankel);
 * periodic, chi_n(y) = exp(2 pi i (n - nmu/2 - 1) y)/sqrt(ny) gives V_mn = D(n-m) with
 * D(q) = (1/ny) sum_i u_i exp(2 pi i q y_i) (Toeplitz, D(-q) = conj D(q) for real u).
 * ------------------------------------------------------------------------------------------- */
int rgo_project(const rgo_wg *w, const double *u, int ustride, zc *V, int method)
{
    int n = w->nmu, ny = w->ny;
    if (method == 0) {
        for (int m = 0; m < n; m++)
            for (int k = 0; k < n; k++) {
                zc s = 0.0;
                for (int i = 0; i < ny; i++)
                    s += conj(w->modal[(size_t)i * n + m]) * (u[(size_t)i * ustride] * w->modal[(size_t)i * n + k]);
                V[(size_t)m * n + k] = s;
            }
        return RGO_OK;
    }
    if (w->boundtype == RGO_DIRICHLET) {
        double *Cq = (double *)calloc((size_t)2 * n + 1, sizeof(double));
        if (!Cq) return RGO_ENOMEM;
        for (int q = 0; q <= 2 * n; q++) {
            double s = 0.0;
            for (int i = 0; i < ny; i++) s += u[(size_t)i * ustride] * cos(q * RGO_PI * w->ymesh[i]);
            Cq[q] = s / ny;
        }
        for (int m = 0; m < n; m++)
            for (int k = 0; k < n; k++)
                V[(size_t)m * n + k] = Cq[abs(m - k)] - Cq[m + k + 2];
        free(Cq);
        return RGO_OK;
    }
    zc *Dq = (zc *)calloc((size_t)n, sizeof(zc));
    if (!Dq) return RGO_ENOMEM;
    for (int q = 0; q < n; q++) {
        zc s = 0.0;
        for (int i = 0; i < ny; i++) s += u[(size_t)i * ustride] * cexp(I * 2.0 * RGO_PI * q * w->ymesh[i]);
        Dq[q] = s / ny;
    }
    for (int m = 0; m < n; m++)
        for (int k = 0; k < n; k++)
            V[(size_t)m * n + k] = k >= m ? Dq[k - m] : conj(Dq[m - k]);
    free(Dq);
    return RGO_OK;
}

void rgo_set_projection(rgo_wg *w, int method) { w->proj_method = method; }

/* A_ix (1-based ix) = diag(a_ix) - V(upot(ix - nxfree, :)): src/waveguide.f90:357-380. */
void rgo_slice_matrix(const rgo_wg *w, int ix, const double *upot, zc *A, zc *work)
{
    int n = w->nmu;
    memset(A, 0, sizeof(zc) * (size_t)n * n);
    for (int m = 0; m < n; m++) {
        zc a;
        if (ix == 1) a = w->abound[m];
        else if (ix == w->nx) a = w->nx == 2 ? w->abound[m] + w->abulk[m] : w->abound[m]; /* :359-360 */
        else a = w->abulk[m];
        A[(size_t)m * n + m] = a;
    }
    int id = ix - w->nxfree;
    if (upot && id >= 1 && id <= w->nxdiso) {
        rgo_project(w, upot + (id - 1), w->nxdiso, work, w->proj_method);
        for (size_t e = 0; e < (size_t)n * n; e++) A[e] -= work[e];
    }
}

/* ---------------------------------------------------------------------------------------------
 * Dense complex kernels.
 * ------------------------------------------------------------------------------------------- */
static inline double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); } /* LAPACK izamax metric */

/* In-place inverse by LU with partial pivoting (zgetrf) followed by inv(U) and the solve
 * inv(A) L = inv(U) with the column interchanges undone (zgetri): src/base_utils.f90:235-274.
 * Returns 0 or RGO_ESINGULAR (the reference stops on LAPACK info /= 0, :255-269). */
int rgo_inverse(int n, zc *A, int *ipiv, zc *work)
{
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = cabs1(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = cabs1(A[(size_t)i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[k] = p;
        if (A[(size_t)p * n + k] == 0.0) return RGO_ESINGULAR;
        if (p != k)
            for (int j = 0; j < n; j++) {
                zc t = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)p * n + j]; A[(size_t)p * n + j] = t;
            }
        zc r = 1.0 / A[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++) {
            zc l = A[(size_t)i * n + k] * r;
            A[(size_t)i * n + k] = l;
            zc *ri = A + (size_t)i * n, *rk = A + (size_t)k * n;
            for (int j = k + 1; j < n; j++) ri[j] -= l * rk[j];
        }
    }
    /* inv(U) in place (upper triangle, non-unit diagonal). */
    for (int j = 0; j < n; j++) {
        A[(size_t)j * n + j] = 1.0 / A[(size_t)j * n + j];
        zc ajj = -A[(size_t)j * n + j];
        /* x = U(0:j-1, j);  x <- invU(0:j-1,0:j-1) * x ; x *= ajj */
        for (int i = 0; i < j; i++) {
            zc s = 0.0;
            for (int k = i; k < j; k++) s += A[(size_t)i * n + k] * A[(size_t)k * n + j];
            work[i] = s;
        }
        for (int i = 0; i < j; i++) A[(size_t)i * n + j] = work[i] * ajj;
    }
    /* Solve X L = invU for X, column by column from the right (unit lower L). */
    for (int j = n - 1; j >= 0; j--) {
        for (int i = j + 1; i < n; i++) { work[i] = A[(size_t)i * n + j]; A[(size_t)i * n + j] = 0.0; }
        for (int i = 0; i < n; i++) {
            zc s = 0.0;
            for (int k = j + 1; k < n; k++) s += A[(size_t)i * n + k] * work[k];
            A[(size_t)i * n + j] -= s;
        }
    }
    for (int j = n - 2; j >= 0; j--) {
        int p = ipiv[j];
        if (p != j)
            for (int i = 0; i < n; i++) {
                zc t = A[(size_t)i * n + j]; A[(size_t)i * n + j] = A[(size_t)i * n + p]; A[(size_t)i * n + p] = t;
            }
    }
    return RGO_OK;
}

int rgo_inverse_alloc(int n, zc *A)
{
    int *ipiv = (int *)malloc(sizeof(int) * n);
    zc *work = (zc *)malloc(sizeof(zc) * n);
    int st = (ipiv && work) ? rgo_inverse(n, A, ipiv, work) : RGO_ENOMEM;
    free(ipiv); free(work);
    return st;
}

/* C = alpha * A(m x k) * B(k x n), row-major, i-k-j order. */
static void rgo_gemm(int m, int n, int k, zc alpha, const zc *A, const zc *B, zc *C)
{
    for (int i = 0; i < m; i++) {
        zc *ci = C + (size_t)i * n;
        for (int j = 0; j < n; j++) ci[j] = 0.0;
        for (int l = 0; l < k; l++) {
            zc a = alpha * A[(size_t)i * k + l];
            const zc *bl = B + (size_t)l * n;
            for (int j = 0; j < n; j++) ci[j] += a * bl[j];
        }
    }
}

/* Singular values of the complex m x n matrix A (row-major), descending, by the one-sided
 * (Hestenes) Jacobi method: plane rotations of column pairs until every pair is orthogonal to
 * working precision; sigma_j = ||a_j||.  Replaces zgesvd(JOBU='N', JOBVT='N') of
 * src/base_utils.f90:341-380 (same output: min(m, n) values in decreasing order).
 * A is not modified.  Returns RGO_ENOCONV if 60 sweeps do not converge. */
int rgo_singular_values(int m, int n, const zc *A, double *S)
{
    int tr = n > m; /* work on A^H when wide so that the column count is min(m, n) */
    int rows = tr ? n : m, cols = tr ? m : n;
    zc *X = (zc *)malloc(sizeof(zc) * (size_t)rows * cols); /* column-major: col j at X + j*rows */
    if (!X) return RGO_ENOMEM;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            if (!tr) X[(size_t)j * rows + i] = A[(size_t)i * n + j];
            else X[(size_t)i * rows + j] = conj(A[(size_t)i * n + j]);
        }
    const double tol = 2.220446049250313e-16 * sqrt((double)rows);
    int conv = 0;
    for (int sweep = 0; sweep < 60 && !conv; sweep++) {
        conv = 1;
        for (int p = 0; p < cols - 1; p++)
            for (int q = p + 1; q < cols; q++) {
                zc *ap = X + (size_t)p * rows, *aq = X + (size_t)q * rows;
                double al = 0.0, be = 0.0;
                zc ga = 0.0;
                for (int i = 0; i < rows; i++) {
                    al += creal(ap[i]) * creal(ap[i]) + cimag(ap[i]) * cimag(ap[i]);
                    be += creal(aq[i]) * creal(aq[i]) + cimag(aq[i]) * cimag(aq[i]);
                    ga += conj(ap[i]) * aq[i];
                }
                double g = cabs(ga);
                if (g == 0.0 || g <= tol * sqrt(al * be)) continue;
                conv = 0;
                double zeta = (be - al) / (2.0 * g);
                double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                zc e = conj(ga) / g; /* exp(-i phi), phi = arg(gamma) */
                for (int i = 0; i < rows; i++) {
                    zc x = ap[i], y = e * aq[i];
                    ap[i] = c * x - s * y;
                    aq[i] = s * x + c * y;
                }
            }
    }
    for (int j = 0; j < cols; j++) {
        double s2 = 0.0;
        for (int i = 0; i < rows; i++) s2 += creal(X[(size_t)j * rows + i]) * creal(X[(size_t)j * rows + i]) +
                                             cimag(X[(size_t)j * rows + i]) * cimag(X[(size_t)j * rows + i]);
        S[j] = sqrt(s2);
    }
    free(X);
    /* insertion sort, descending */
    for (int i = 1; i < cols; i++) {
        double v = S[i]; int k = i - 1;
        while (k >= 0 && S[k] < v) { S[k + 1] = S[k]; k--; }
        S[k + 1] = v;
    }
    return conv ? RGO_OK : RGO_ENOCONV;
}

/* ---------------------------------------------------------------------------------------------
 * RGF sweep: src/waveguide.f90:427-440.  On return G1 = Gs1 = G(L; 0) (nmu x nmu).
 * ------------------------------------------------------------------------------------------- */
typedef struct { zc *Gss, *Gs1, *A, *T1, *T2, *work; int *ipiv; } rgo_ws;

static int rgo_ws_alloc(rgo_ws *s, int n)
{
    size_t nn = (size_t)n * n;
    s->Gss = (zc *)malloc(sizeof(zc) * nn); s->Gs1 = (zc *)malloc(sizeof(zc) * nn);
    s->A = (zc *)malloc(sizeof(zc) * nn); s->T1 = (zc *)malloc(sizeof(zc) * nn);
    s->T2 = (zc *)malloc(sizeof(zc) * nn); s->work = (zc *)malloc(sizeof(zc) * nn);
    s->ipiv = (int *)malloc(sizeof(int) * n);
    return (s->Gss && s->Gs1 && s->A && s->T1 && s->T2 && s->work && s->ipiv) ? RGO_OK : RGO_ENOMEM;
}
static void rgo_ws_free(rgo_ws *s)
{
    free(s->Gss); free(s->Gs1); free(s->A); free(s->T1); free(s->T2); free(s->work); free(s->ipiv);
}

static int rgo_sweep(const rgo_wg *w, const double *upot, rgo_ws *s)
{
    int n = w->nmu;
    size_t nn = (size_t)n * n;
    rgo_slice_matrix(w, 1, upot, s->Gss, s->work);                          /* :427 */
    int st = rgo_inverse(n, s->Gss, s->ipiv, s->work);                       /* :428 */
    if (st) return st;
    memcpy(s->Gs1, s->Gss, sizeof(zc) * nn);                                 /* :429 */
    for (int ix = 2; ix <= w->nx; ix++) {                                    /* :431 */
        rgo_slice_matrix(w, ix, upot, s->A, s->work);
        for (int m = 0; m < n; m++)                                          /* :434 */
            for (int k = 0; k < n; k++)
                s->Gss[(size_t)m * n + k] = s->A[(size_t)m * n + k] - (w->bmat[m] * s->Gss[(size_t)m * n + k]) * w->bmat[k];
        st = rgo_inverse(n, s->Gss, s->ipiv, s->work);                       /* :435 */
        if (st) return st;
        for (int m = 0; m < n; m++)                                          /* :438: B * Gs1 */
            for (int k = 0; k < n; k++) s->T1[(size_t)m * n + k] = w->bmat[m] * s->Gs1[(size_t)m * n + k];
        rgo_gemm(n, n, n, -1.0, s->Gss, s->T1, s->Gs1);
    }
    return RGO_OK;
}

/* t_ij = 2 i sqrt(mu_i) G_ij sqrt(mu_j) over the filtered ranges: src/waveguide.f90:443-445.
 * t is (imax-imin+1) x (jmax-jmin+1), row-major. */
static void rgo_fisher_lee(const rgo_wg *w, const zc *G, zc *t)
{
    int n = w->nmu, nr = w->imax - w->imin + 1, nc = w->jmax - w->jmin + 1;
    for (int a = 0; a < nr; a++)
        for (int b = 0; b < nc; b++) {
            int i = w->imin - 1 + a, j = w->jmin - 1 + b;
            t[(size_t)a * nc + b] = 2.0 * I * ((csqrt(w->mumesh[i]) * G[(size_t)i * n + j]) * csqrt(w->mumesh[j]));
        }
}

/* Transmission matrix of one realization with explicit potential (NULL = no disorder). */
int rgo_transmission_matrix(const rgo_wg *w, const double *upot, zc *t)
{
    rgo_ws s;
    int st = rgo_ws_alloc(&s, w->nmu);
    if (!st) st = rgo_sweep(w, upot, &s);
    if (!st) rgo_fisher_lee(w, s.Gs1, t);
    rgo_ws_free(&s);
    return st;
}

/* G(L;0) = Gs1 after the sweep (nmu x nmu). */
int rgo_green_l0(const rgo_wg *w, const double *upot, zc *G)
{
    rgo_ws s;
    int st = rgo_ws_alloc(&s, w->nmu);
    if (!st) st = rgo_sweep(w, upot, &s);
    if (!st) memcpy(G, s.Gs1, sizeof(zc) * (size_t)w->nmu * w->nmu);
    rgo_ws_free(&s);
    return st;
}

/* Transmission eigenvalues T = s^2 (descending) of one realization: src/waveguide.f90:392-462. */
int rgo_tdistrib_potential(const rgo_wg *w, const double *upot, double *teig)
{
    int nr = w->imax - w->imin + 1, nc = w->jmax - w->jmin + 1;
    zc *t = (zc *)malloc(sizeof(zc) * (size_t)nr * nc);
    double *S = (double *)malloc(sizeof(double) * (nr < nc ? nr : nc));
    if (!t || !S) { free(t); free(S); return RGO_ENOMEM; }
    int st = rgo_transmission_matrix(w, upot, t);
    if (!st) st = rgo_singular_values(nr, nc, t, S);                         /* :448 */
    if (!st || st == RGO_ENOCONV)
        for (int k = 0; k < w->neigv; k++) teig[k] = S[k] * S[k];            /* :449 */
    free(t); free(S);
    return st;
}

int rgo_tdistrib_seed(const rgo_wg *w, uint64_t seed, double *teig)
{
    double *u = (double *)malloc(sizeof(double) * (size_t)w->nxdiso * w->ny);
    if (!u) return RGO_ENOMEM;
    rgo_potential(w, seed, u);
    int st = rgo_tdistrib_potential(w, u, teig);
    free(u);
    return st;
}

/* OpenMP loop over realizations (src/main.f90:148-171): seeds seed0 .. seed0+nseed-1, chunk k of
 * tallvals (neigv values) = realization k.  Returns the first non-zero status. */
int rgo_tdistrib_batch(const rgo_wg *w, uint64_t seed0, int nseed, int nthreads, double *tallvals)
{
    int status = RGO_OK;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int k = 0; k < nseed; k++) {
        int st = rgo_tdistrib_seed(w, seed0 + (uint64_t)k, tallvals + (size_t)k * w->neigv);
        if (st) {
#ifdef _OPENMP
#pragma omp critical
#endif
            if (!status) status = st;
        }
    }
    return status;
}

/* ---------------------------------------------------------------------------------------------
 * r and t, version 1 (src/waveguide.f90:469-519): the sweep also carries G11 and G1s;
 * Fisher-Lee with sqrt(Re mu) (:510-512).  r, t: nmu x nmu row-major.
 * ------------------------------------------------------------------------------------------- */
int rgo_refl_trans_v1(const rgo_wg *w, const double *upot, zc *r, zc *t)
{
    int n = w->nmu;
    size_t nn = (size_t)n * n;
    rgo_ws s;
    int st = rgo_ws_alloc(&s, n);
    zc *G1s = (zc *)malloc(sizeof(zc) * nn), *G11 = (zc *)malloc(sizeof(zc) * nn);
    if (st || !G1s || !G11) { rgo_ws_free(&s); free(G1s); free(G11); return RGO_ENOMEM; }
    rgo_slice_matrix(w, 1, upot, s.Gss, s.work);
    st = rgo_inverse(n, s.Gss, s.ipiv, s.work);
    if (!st) {
        memcpy(s.Gs1, s.Gss, sizeof(zc) * nn); memcpy(G1s, s.Gss, sizeof(zc) * nn); memcpy(G11, s.Gss, sizeof(zc) * nn);
        for (int ix = 2; ix <= w->nx && !st; ix++) {
            rgo_slice_matrix(w, ix, upot, s.A, s.work);
            for (int m = 0; m < n; m++)
                for (int k = 0; k < n; k++)
                    s.Gss[(size_t)m * n + k] = s.A[(size_t)m * n + k] - (w->bmat[m] * s.Gss[(size_t)m * n + k]) * w->bmat[k];
            st = rgo_inverse(n, s.Gss, s.ipiv, s.work);                       /* :496 */
            if (st) break;
            for (int m = 0; m < n; m++)
                for (int k = 0; k < n; k++) s.T1[(size_t)m * n + k] = w->bmat[m] * s.Gs1[(size_t)m * n + k];
            rgo_gemm(n, n, n, -1.0, s.Gss, s.T1, s.Gs1);                     /* :499 */
            for (int m = 0; m < n; m++)
                for (int k = 0; k < n; k++) s.T1[(size_t)m * n + k] = w->bmat[m] * s.Gs1[(size_t)m * n + k];
            rgo_gemm(n, n, n, 1.0, G1s, s.T1, s.T2);                          /* :502 */
            for (size_t e = 0; e < nn; e++) G11[e] -= s.T2[e];
            for (int m = 0; m < n; m++)
                for (int k = 0; k < n; k++) s.T1[(size_t)m * n + k] = w->bmat[m] * s.Gss[(size_t)m * n + k];
            rgo_gemm(n, n, n, -1.0, G1s, s.T1, s.T2);                         /* :505 */
            memcpy(G1s, s.T2, sizeof(zc) * nn);
        }
    }
    if (!st)
        for (int i = 0; i < n; i++) {
            double si = sqrt(creal(w->mumesh[i]));
            for (int j = 0; j < n; j++) {
                double sj = sqrt(creal(w->mumesh[j]));
                t[(size_t)i * n + j] = 2.0 * I * ((si * s.Gs1[(size_t)i * n + j]) * sj);
                r[(size_t)i * n + j] = (i == j ? -1.0 : 0.0) + 2.0 * I * ((si * G11[(size_t)i * n + j]) * sj);
            }
        }
    rgo_ws_free(&s); free(G1s); free(G11);
    return st;
}

/* Block-tridiagonal solve H X = Imat, H = tridiag(B, A_i, B), B = diag(bdiag):
 * src/base_utils.f90:296-335.  amat: nb blocks of bs x bs; imat, xmat: nb blocks of bs x nrhs. */
int rgo_solve_block_tridiag(int nb, int bs, int nrhs, const zc *amat, const zc *bdiag, const zc *imat, zc *xmat)
{
    size_t bb = (size_t)bs * bs, br = (size_t)bs * nrhs;
    zc *smat = (zc *)malloc(sizeof(zc) * bb * nb), *zmat = (zc *)malloc(sizeof(zc) * br * nb);
    zc *tmp = (zc *)malloc(sizeof(zc) * (bb > br ? bb : br)), *work = (zc *)malloc(sizeof(zc) * bs);
    int *ipiv = (int *)malloc(sizeof(int) * bs);
    int st = (smat && zmat && tmp && work && ipiv) ? RGO_OK : RGO_ENOMEM;
    if (!st) {
        memcpy(smat, amat, sizeof(zc) * bb);                                     /* :320 */
        st = rgo_inverse(bs, smat, ipiv, work);                                 /* :321 */
        if (!st) rgo_gemm(bs, nrhs, bs, 1.0, smat, imat, zmat);                 /* :322 */
        for (int i = 1; i < nb && !st; i++) {                                    /* :324 */
            zc *S = smat + bb * i, *Sp = smat + bb * (i - 1);
            for (int m = 0; m < bs; m++)
                for (int k = 0; k < bs; k++)
                    S[(size_t)m * bs + k] = amat[bb * i + (size_t)m * bs + k] - (bdiag[m] * Sp[(size_t)m * bs + k]) * bdiag[k];
            st = rgo_inverse(bs, S, ipiv, work);                                 /* :326 */
            if (st) break;
            for (int m = 0; m < bs; m++)                                         /* :327 */
                for (int k = 0; k < nrhs; k++)
                    tmp[(size_t)m * nrhs + k] = imat[br * i + (size_t)m * nrhs + k] - bdiag[m] * zmat[br * (i - 1) + (size_t)m * nrhs + k];
            rgo_gemm(bs, nrhs, bs, 1.0, S, tmp, zmat + br * i);
        }
        if (!st) {
            memcpy(xmat + br * (nb - 1), zmat + br * (nb - 1), sizeof(zc) * br); /* :330 */
            for (int i = nb - 2; i >= 0; i--) {                                  /* :331-333 */
                for (int m = 0; m < bs; m++)
                    for (int k = 0; k < nrhs; k++) tmp[(size_t)m * nrhs + k] = bdiag[m] * xmat[br * (i + 1) + (size_t)m * nrhs + k];
                rgo_gemm(bs, nrhs, bs, 1.0, smat + bb * i, tmp, xmat + br * i);
                for (size_t e = 0; e < br; e++) xmat[br * i + e] = zmat[br * i + e] - xmat[br * i + e];
            }
        }
    }
    free(smat); free(zmat); free(tmp); free(work); free(ipiv);
    return st;
}

/* r and t, version 2 (src/waveguide.f90:526-553). */
int rgo_refl_trans_v2(const rgo_wg *w, const double *upot, zc *r, zc *t)
{
    int n = w->nmu, nx = w->nx;
    size_t nn = (size_t)n * n;
    zc *amat = (zc *)malloc(sizeof(zc) * nn * nx), *ind = (zc *)calloc(nn * nx, sizeof(zc));
    zc *green = (zc *)malloc(sizeof(zc) * nn * nx), *work = (zc *)malloc(sizeof(zc) * nn);
    int st = (amat && ind && green && work) ? RGO_OK : RGO_ENOMEM;
    if (!st) {
        for (int ix = 1; ix <= nx; ix++) rgo_slice_matrix(w, ix, upot, amat + nn * (ix - 1), work);
        for (int m = 0; m < n; m++) ind[(size_t)m * n + m] = 1.0;                /* :540 */
        st = rgo_solve_block_tridiag(nx, n, n, amat, w->bmat, ind, green);
    }
    if (!st)
        for (int i = 0; i < n; i++) {
            double si = sqrt(creal(w->mumesh[i]));
            for (int j = 0; j < n; j++) {
                double sj = sqrt(creal(w->mumesh[j]));
                t[(size_t)i * n + j] = 2.0 * I * ((si * green[nn * (nx - 1) + (size_t)i * n + j]) * sj); /* :547 */
                r[(size_t)i * n + j] = (i == j ? -1.0 : 0.0) + 2.0 * I * ((si * green[(size_t)i * n + j]) * sj); /* :548 */
            }
        }
    free(amat); free(ind); free(green); free(work);
    return st;
}

/* Deposition matrix (src/waveguide.f90:710-745): the column block ipos of H^-1 by the block-tridiagonal
 * solve with the identity at block ipos (:736-739), then dep = 2i s G(fpos; ipos) s with s = sqrt(Re mu)
 * (:741-742).  fpos, ipos in 1..nx. */
int rgo_deposition(const rgo_wg *w, const double *upot, int fpos, int ipos, zc *dep)
{
    int n = w->nmu, nx = w->nx;
    if (fpos < 1 || fpos > nx || ipos < 1 || ipos > nx) return RGO_EINVAL;
    size_t nn = (size_t)n * n;
    zc *amat = (zc *)malloc(sizeof(zc) * nn * nx), *ind = (zc *)calloc(nn * nx, sizeof(zc));
    zc *green = (zc *)malloc(sizeof(zc) * nn * nx), *work = (zc *)malloc(sizeof(zc) * nn);
    int st = (amat && ind && green && work) ? RGO_OK : RGO_ENOMEM;
    if (!st) {
        for (int ix = 1; ix <= nx; ix++) rgo_slice_matrix(w, ix, upot, amat + nn * (ix - 1), work);
        for (int m = 0; m < n; m++) ind[nn * (ipos - 1) + (size_t)m * n + m] = 1.0;
        st = rgo_solve_block_tridiag(nx, n, n, amat, w->bmat, ind, green);
    }
    if (!st)
        for (int i = 0; i < n; i++) {
            double si = sqrt(creal(w->mumesh[i]));
            for (int j = 0; j < n; j++) {
                double sj = sqrt(creal(w->mumesh[j]));
                dep[(size_t)i * n + j] = 2.0 * I * ((si * green[nn * (fpos - 1) + (size_t)i * n + j]) * sj);
            }
        }
    free(amat); free(ind); free(green); free(work);
    return st;
}

/* Column block ipos of H^-1 for every slice, green[x] = G(x; ipos) (nx blocks of nmu x nmu, row-major),
 * by the block-tridiagonal solve of compute_tprofile (src/waveguide.f90:873-878). */
int rgo_green_column(const rgo_wg *w, const double *upot, int ipos, zc *green)
{
    int n = w->nmu, nx = w->nx;
    if (ipos < 1 || ipos > nx) return RGO_EINVAL;
    size_t nn = (size_t)n * n;
    zc *amat = (zc *)malloc(sizeof(zc) * nn * nx), *ind = (zc *)calloc(nn * nx, sizeof(zc));
    zc *work = (zc *)malloc(sizeof(zc) * nn);
    int st = (amat && ind && work) ? RGO_OK : RGO_ENOMEM;
    if (!st) {
        for (int ix = 1; ix <= nx; ix++) rgo_slice_matrix(w, ix, upot, amat + nn * (ix - 1), work);
        for (int m = 0; m < n; m++) ind[nn * (ipos - 1) + (size_t)m * n + m] = 1.0;
        st = rgo_solve_block_tridiag(nx, n, n, amat, w->bmat, ind, green);
    }
    free(amat); free(ind); free(work);
    return st;
}

/* ---------------------------------------------------------------------------------------------
 * Histogram chain: src/base_utils.f90:491-570, src/main.f90:223-251, src/plot_histogram.f90:80-83.
 * ------------------------------------------------------------------------------------------- */
/* bintype 0: linear  xmin + (xmax-xmin)*(i-1)/nbins ; 1: Chebyshev xmin + (xmax-xmin)*(1-cos(pi*(i-1)/nbins))/2 */
void rgo_bins(int bintype, int nbins, double xmin, double xmax, double *bins)
{
    for (int i = 0; i <= nbins; i++) {
        if (bintype == 0) bins[i] = xmin + (xmax - xmin) * i / nbins;
        else bins[i] = xmin + (xmax - xmin) * (1.0 - cos(RGO_PI * i / nbins)) / 2;
    }
}

/* Counts with the last edge inclusive; out-of-range samples are "missing"; NaN falls through
 * every comparison and is dropped silently (src/base_utils.f90:501-514). */
void rgo_histogram(int64_t nsamp, const double *samples, int nbins, const double *bins, double *histo, double *missing)
{
    *missing = 0.0;
    for (int j = 0; j < nbins; j++) histo[j] = 0.0;
    for (int64_t i = 0; i < nsamp; i++) {
        double x = samples[i];
        if (x < bins[0] || bins[nbins] < x) *missing += 1.0;
        else if (x == bins[nbins]) histo[nbins - 1] += 1.0;
        else
            for (int j = 0; j < nbins; j++)
                if (x < bins[j + 1]) { histo[j] += 1.0; break; }
    }
}

void rgo_normalize_histogram(int nbins, const double *bins, double *histo)
{
    double npoints = 0.0;
    for (int i = 0; i < nbins; i++) npoints += histo[i];
    for (int i = 0; i < nbins; i++) histo[i] = histo[i] / (npoints * (bins[i + 1] - bins[i]));
}

void rgo_moving_average(int nsig, const double *signal, int n, double *res)
{
    for (int i = 0; i + n <= nsig; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += signal[i + j];
        res[i] = s / n;
    }
}

/* prepare_distrib (src/main.f90:223-251) generalised to either bin type on [xmin, xmax];
 * rho_data: nbins x 2 row-major (tm, rho).  Returns the number of missing samples. */
double rgo_prepare_distrib(int64_t nsamp, const double *samples, int nbins, int bintype, double xmin, double xmax,
                           double *rho_data)
{
    double *bins = (double *)malloc(sizeof(double) * (nbins + 1));
    double *histo = (double *)malloc(sizeof(double) * nbins);
    double *cent = (double *)malloc(sizeof(double) * nbins);
    double missing = 0.0;
    rgo_bins(bintype, nbins, xmin, xmax, bins);
    rgo_histogram(nsamp, samples, nbins, bins, histo, &missing);
    for (int i = 0; i < nbins; i++) histo[i] = histo[i] + 2.220446049250313e-16 / 10.0; /* :243 */
    rgo_normalize_histogram(nbins, bins, histo);
    rgo_moving_average(nbins + 1, bins, 2, cent);
    for (int i = 0; i < nbins; i++) { rho_data[2 * i] = cent[i]; rho_data[2 * i + 1] = histo[i]; }
    free(bins); free(histo); free(cent);
    return missing;
}

int rgo_omp_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double rgo_wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}
