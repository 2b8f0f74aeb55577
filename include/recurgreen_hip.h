/*
 * recurgreen_hip.h -- C ABI of the MI355X (gfx950) recursive Green function library.
 *
 * Drop-in for the CPU/LAPACK transmission path of RecurGreen's `waveguide` module.  The
 * reference has no FFI layer; its boundary is the Fortran type `waveguide_t` and its
 * type-bound procedures (src/waveguide.f90:13-73).  Each entry point below names the
 * reference procedure it replaces.  The Fortran binding is recurgreen_amd/fortran/
 * recurgreen_hip.f90 (iso_c_binding); the Python binding is recurgreen_amd/binding.py
 * (ctypes).
 *
 * Conventions
 *   - Plain C types only.  Complex numbers are interleaved (re, im) doubles, the layout of
 *     Fortran complex(real64) and C double _Complex.
 *   - Matrices crossing the ABI are column-major (Fortran order): element (i, j) of an
 *     m x n matrix at [2*(i + j*m)] / [2*(i + j*m) + 1].
 *   - A potential batch is upot(ix, iy, r): ix fastest (src/waveguide.f90:43), in units of
 *     U*dx/k, shape (nxdiso, ny, nreal).
 *   - Transmission eigenvalues: realization r occupies tallvals[r*neigv .. r*neigv+neigv-1],
 *     in decreasing order (src/waveguide.f90:448-454).
 *   - Every function returns an rg_status; 0 is success.  Pointers documented as "host" are
 *     ordinary host memory; "device" pointers are HIP device allocations on the plan's
 *     device.  A plan is used by one caller at a time; plans are independent.
 */
#ifndef RECURGREEN_HIP_H
#define RECURGREEN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rg_status {
    RG_OK = 0,
    RG_EINVAL = 1,       /* invalid argument; the reference stops with errmsg_invalid_arg */
    RG_ESINGULAR = 2,    /* a slice matrix was numerically singular (zgetrf info > 0) */
    RG_ENOCONV = 3,      /* the singular-value iteration did not converge (zgesvd info > 0) */
    RG_ENOMEM = 4,       /* device or host allocation failed */
    RG_EDEVICE = 5,      /* HIP runtime error */
    RG_ENOTSUP = 6,      /* combination not supported by this build (see rg_status_string) */
    RG_ENONFINITE = 7    /* non-finite value in an input potential or in G(L;0) */
} rg_status;

typedef enum rg_boundtype { RG_DIRICHLET = 0, RG_PERIODIC = 1 } rg_boundtype;

typedef enum rg_precision {
    RG_FP64 = 0,         /* every step in complex128 */
    RG_MIXED_FP32 = 1,   /* slice GEMMs and inversion updates in FP32, FP64 pivot decisions */
    RG_MIXED_BF16X3 = 2  /* GEMMs as 3-term split-BF16 products with FP32 accumulation */
} rg_precision;

/* Parameters of waveguide_t%init (src/waveguide.f90:81), plus two extensions:
 *   ny        > 0 sets the transverse grid (config mode, ny >= nmu); 0 keeps ny = nmu
 *             (src/waveguide.f90:162,202).
 *   dos_model 0 = the reference's finite mode-sum DOS (:243); 1 = continuum 1/(4 pi). */
typedef struct rg_params {
    int32_t boundtype;
    int32_t nxdiso, nxfree;
    double wol, exfac, aratio, dscat, dabso, napera, naperb;
    int32_t ny;
    int32_t dos_model;
} rg_params;

typedef struct rg_plan rg_plan;

/* Mirrors waveguide_t%init (src/waveguide.f90:81-255) and get_nmu/get_neigv/get_nx (:291-314).
 * batch_max is the number of realizations one device pass holds (device memory is allocated
 * here, never per call). */
rg_status rg_plan_create(const rg_params *params, int32_t batch_max, int32_t device, rg_plan **plan);
/* Mirrors waveguide_t%del (src/waveguide.f90:275-286). */
rg_status rg_plan_destroy(rg_plan *plan);
/* dims[0..10] = nmu, ny, nx, neigv, imin, imax, jmin, jmax, nprop, nxdiso, nxfree (1-based
 * index ranges like the reference).  vals[0..3] = kL, k*dx, dos, stdu (may be NULL). */
rg_status rg_plan_dims(const rg_plan *plan, int32_t *dims, double *vals);

/* Mirrors compute_tdistrib (src/waveguide.f90:392-462) over a batch of seeds, the loop of
 * src/main.f90:148-171: realization r uses seed seed0 + r; tallvals (host) receives
 * nseed*neigv values, status (host, may be NULL) one rg_status per realization.  The disorder
 * is generated on the device (Philox4x32-10 keyed by seed, ix, iy / 2; both outputs of the
 * Box-Muller transform of src/random_normal.f90:64, cosine for even iy, sine for odd iy). */
rg_status rg_tdistrib(rg_plan *plan, int64_t seed0, int32_t nseed, int32_t precision,
                      double *tallvals, int32_t *status);
/* Same with device-resident output (tallvals_dev: nseed*neigv doubles, status_dev: nseed
 * int32, may be NULL); the call is asynchronous on the plan's stream. */
rg_status rg_tdistrib_device(rg_plan *plan, int64_t seed0, int32_t nseed, int32_t precision,
                             double *tallvals_dev, int32_t *status_dev);
/* Uneven batch (no counterpart in the reference, whose OpenMP loop runs identical copies of one
 * waveguide_t): realization r has its own parameters params[r] and seed seeds[r] -- any length,
 * absorption and disorder strength, and any width.  Realizations sharing boundary type, mode count,
 * transverse cells, filter ranges and free length are swept together (sorted by length, the
 * finished ones dropping out); other widths form further groups.  Output in the caller's order:
 * the eigenvalues of realization r (descending) at tallvals[offsets[r] .. offsets[r+1]), offsets
 * has nreal+1 entries; status (may be NULL) one rg_status per realization (RG_ENOTSUP for mode
 * counts above 256).  Allocates one plan per group of at most batch_max realizations. */
rg_status rg_tdistrib_uneven(const rg_params *params, const int64_t *seeds, int32_t nreal, int32_t precision,
                             int32_t device, int32_t batch_max, double *tallvals, int64_t *offsets, int32_t *status);
/* Explicit potentials (host upot, nxdiso*ny*nreal doubles): the path used for parity against
 * a CPU implementation fed the same potentials. */
rg_status rg_tdistrib_potential(rg_plan *plan, const double *upot, int32_t nreal, int32_t precision,
                                double *tallvals, int32_t *status);
/* Transmission matrices t (filtered ranges, (imax-imin+1) x (jmax-jmin+1) column-major
 * complex each) for explicit host potentials; upot == NULL means no disorder. */
rg_status rg_transmission_potential(rg_plan *plan, const double *upot, int32_t nreal, int32_t precision,
                                    double *tmat, int32_t *status);
/* Mirrors compute_refl_trans_v1 (src/waveguide.f90:469-519) for seeds seed0 .. seed0+nseed-1:
 * the reflection and transmission matrices r = -I + 2i s G11 s, t = 2i s G(L;0) s with
 * s = sqrt(Re mu) (:508-512), full nmu x nmu, column-major complex, realization-major in rmat
 * and tmat (host, nseed*nmu*nmu complex each).  Complex128 only (the sweep of :484-506 with G11
 * and G1s carried beside Gss and Gs1). */
rg_status rg_refl_trans(rg_plan *plan, int64_t seed0, int32_t nseed, double *rmat, double *tmat, int32_t *status);
/* Same for explicit host potentials upot (nxdiso*ny*nreal doubles). */
rg_status rg_refl_trans_potential(rg_plan *plan, const double *upot, int32_t nreal, double *rmat, double *tmat,
                                  int32_t *status);
/* Mirrors compute_deposition (src/waveguide.f90:710-745) for seeds seed0 .. seed0+nseed-1: the
 * deposition matrix dep = 2i s G(fpos; ipos) s, s = sqrt(Re mu) (:741-742), full nmu x nmu,
 * column-major complex, realization-major in depmat (host, nseed*nmu*nmu complex).  fpos, ipos in
 * 1..nx (RG_EINVAL otherwise).  G(fpos; ipos) comes from a sweep from the left lead up to fpos-1 and
 * one from the right lead down to fpos+1, joined at slice fpos; the sweep on the side of ipos also
 * carries the propagator from slice ipos.  Complex128 only. */
rg_status rg_deposition(rg_plan *plan, int64_t seed0, int32_t nseed, int32_t fpos, int32_t ipos, double *depmat,
                        int32_t *status);
/* Same for explicit host potentials upot (nxdiso*ny*nreal doubles). */
rg_status rg_deposition_potential(rg_plan *plan, const double *upot, int32_t nreal, int32_t fpos, int32_t ipos,
                                  double *depmat, int32_t *status);
/* Transmission eigenstates inside the medium, the data of compute_tprofile and
 * compute_tprofile_group (src/waveguide.f90:843-997) for seeds seed0 .. seed0+nseed-1, complex128.
 * t = 2i s G(L;0) s, s = sqrt(Re mu): the full nmu x nmu matrix (compute_tprofile, filtered = 0) or
 * rows imin..imax and columns jmin..jmax (compute_tprofile_group, filtered = 1; RG_EINVAL if it has
 * more columns than rows, as diagonalize_ahca's size check).  With nv its number of columns, tvals
 * (host, nseed*nv) receives the eigenvalues T_v of t^H t, descending; intensity (host, nseed*nx*nv,
 * index (r*nx + x)*nv + v) receives |G(x;1) s U(:,v)|^2 summed over the modes, U the eigenvectors
 * (diagonalize_ahca, src/base_utils.f90:394-443).  The reference's profiles are sums of these over
 * the eigenvalues in a window (:858-869, :960-985).  rg_tprofile_reserve allocates the workspace
 * (the Green functions of every slice for nreal_max realizations per device pass; the only
 * allocating call of this group): rg_tprofile returns RG_EINVAL before it, RG_ENOTSUP above nmu = 256. */
rg_status rg_tprofile_reserve(rg_plan *plan, int32_t nreal_max);
rg_status rg_tprofile(rg_plan *plan, int64_t seed0, int32_t nseed, int32_t filtered, double *tvals, double *intensity,
                      int32_t *status);
/* Same for explicit host potentials upot (nxdiso*ny*nreal doubles). */
rg_status rg_tprofile_potential(rg_plan *plan, const double *upot, int32_t nreal, int32_t filtered, double *tvals,
                                double *intensity, int32_t *status);
/* Mirrors compute_wavefunction (src/waveguide.f90:636-697) for the realization of `seed`: psi (host,
 * nx*ny complex, index ix*ny + iy, 0-based) receives psi(x_ix, y_iy) = sum_m chi_m(y_iy) G(x_ix; ipos)[m,
 * imu] for the incident mode imode (1..nmu; imode = 1 is the frontal mode, :657-661) and the source
 * slice ipos (1..nx), RG_EINVAL otherwise (:646-654).  The slice Green functions are kept in the
 * tprofile workspace: call rg_tprofile_reserve(plan, 1) or more first.  Complex128. */
rg_status rg_wavefunction(rg_plan *plan, int64_t seed, int32_t ipos, int32_t imode, double *psi, int32_t *status);
/* Same for an explicit host potential upot (nxdiso*ny doubles). */
rg_status rg_wavefunction_potential(rg_plan *plan, const double *upot, int32_t ipos, int32_t imode, double *psi,
                                    int32_t *status);
/* Adjoint sweep (no counterpart in the reference: the reverse-mode derivative of the sweep of
 * src/waveguide.f90:427-445).  loss_kind 0: L = Tr(t^H t) = sum of the transmission eigenvalues;
 * loss_kind 1: L = T_max, the largest transmission eigenvalue (t over the filtered ranges, complex
 * sqrt(mu) as in :443; the derivative is that of a simple eigenvalue; its eigenvector comes from the
 * eigensolver up to order 128 and from a one-sided Jacobi decomposition up to 256).  For each realization: loss[r] = L and
 * grad = dL/d upot with the potential layout (nxdiso, ny, nreal), units of upot (U dx / k; dL/dU =
 * grad * dx / k).  Complex128; slice Green functions are recomputed from checkpoints kept every
 * ceil(sqrt(nx)) slices.  rg_adjoint_reserve allocates the workspace for nreal_max realizations
 * per device pass (the only allocating adjoint call); the adjoint calls return RG_EINVAL before. */
rg_status rg_adjoint_reserve(rg_plan *plan, int32_t nreal_max);
/* Host buffers: upot (nxdiso*ny*nreal), loss (nreal), grad (nxdiso*ny*nreal), status (may be NULL). */
rg_status rg_adjoint_potential(rg_plan *plan, const double *upot, int32_t nreal, int32_t loss_kind, double *loss,
                               double *grad, int32_t *status);
/* Device buffers, asynchronous on the plan's stream (status_dev may be NULL). */
rg_status rg_adjoint_potential_device(rg_plan *plan, const double *upot_dev, int32_t nreal, int32_t loss_kind,
                                      double *loss_dev, double *grad_dev, int32_t *status_dev);
/* Disorder potentials the device generator produces for seeds seed0.. (host upot out,
 * nxdiso*ny*nreal doubles): lets a caller feed the identical disorder to another code. */
rg_status rg_generate_potential(rg_plan *plan, int64_t seed0, int32_t nreal, double *upot);

/* Histogram of samples (device or host pointer, flag `on_device`) on bins[0..nbins] (host),
 * with the counting rules of src/base_utils.f90:491-516: out-of-range samples are counted in
 * *missing, the last edge is inclusive, NaN is dropped and counted in *dropped.
 * counts (host) receives nbins int64 values. */
rg_status rg_histogram(rg_plan *plan, const double *samples, int64_t nsamp, int32_t on_device,
                       const double *bins, int32_t nbins, int64_t *counts, int64_t *missing,
                       int64_t *dropped);

/* No counterpart in the reference: rg_histogram for device-resident samples that allocates nothing
 * (no plan, no hipMalloc/hipFree, no host buffers), for callers that cannot allocate -- code that
 * runs between a graph capture's set-up and its launch, serving loops, harnesses that forbid
 * allocation.  samples_dev: nsamp doubles on GPU `device`; bins (host): nbins + 1 sorted edges,
 * 1 <= nbins <= 4096 (RG_ENOTSUP above); stream: a hipStream_t on that device (NULL: the null
 * stream).  Same counting rules and outputs (host) as rg_histogram; returns after the stream has
 * reached the end of the call.  Scratch: a static 64 KiB device array of the library, one per GPU;
 * calls that overlap on one GPU race on it. */
rg_status rg_histogram_noalloc(int32_t device, void *stream, const double *samples_dev, int64_t nsamp,
                               const double *bins, int32_t nbins, int64_t *counts, int64_t *missing,
                               int64_t *dropped);

/* Stream the plan launches on (a hipStream_t), for callers that order their own work. */
rg_status rg_plan_stream(rg_plan *plan, void **stream);
rg_status rg_plan_synchronize(rg_plan *plan);
/* Per-kernel timing with HIP events on the plan's stream.  enable != 0 starts recording;
 * rg_plan_kernel_times fills up to `cap` entries of names (NUL-separated, 32 chars each),
 * total milliseconds and launch counts since the last reset. */
rg_status rg_plan_profile(rg_plan *plan, int32_t enable);
rg_status rg_plan_kernel_times(rg_plan *plan, int32_t cap, char *names, double *ms, int64_t *launches,
                               int32_t *count);

/* Launch options (no counterpart in the reference: the choices the autotuner makes, module
 * recurgreen_amd.autotune).  Keys and values:
 *   "inverse_fp64"  0 register-resident Gauss-Jordan, 1 look-ahead Gauss-Jordan, 2 blocked LU, 3 verified
 *                   diagonal pivots on the FP64 matrix cores (pivoted look-ahead for the rest), 4 (default)
 *                   the same on the lower triangle of the complex symmetric slice matrices of Dirichlet
 *                   guides (block sweep operator; 3 for periodic guides);
 *   "inverse_mixed" 1 look-ahead, 2 blocked, 3 verified diagonal pivots, 4 (default) the same on the
 *                   lower triangle of complex symmetric slice matrices (Dirichlet; 3 for periodic guides);
 *   "dyson"         complex64 Dyson product: 0 64x64 output tiles, 1 one workgroup per realization
 *                   (NP 80..128), 2 the same on the BF16 matrix cores with every FP32
 *                   operand split into three BF16 values (six products, FP32 accuracy per
 *                   product; NP 96 and 128, else 1; faster, but the N = 128 sweep's mean T moved
 *                   4.7e-7 from complex128 against 2.0e-8 with 1, so not the default), 3 (default)
 *                   one workgroup per realization with three real products per complex product
 *                   (Gauss's form; NP 96 and 128 whole, 160, 192, 224, 256 as
 *                   2 x 2 blocks, else 1);
 *   "dyson_fp64"    complex128 Dyson product: 0 64x64 output tiles, 1 (default) one workgroup per
 *                   realization (NP 80..128) or per quadrant (NP 160, 192, 224, 256), 2 the same
 *                   with three real products per complex product (Gauss's form);
 *   "project_parts" 1..8 partial sums of the projection over transverse cells;
 *   "project_wide_min" >= 1: batches of at least this many realizations (default 1: all) project
 *                   on the wide kernels (128 or 64 realizations per workgroup, the table staged
 *                   straight into LDS, seeded disorder generated in the kernel up to 128 modes);
 *                   smaller ones on the 32-realization kernel after a separate noise kernel;
 *   "refine"        0..8 (default 0): complex64 sweeps of uniform batches refine G(L;0) in FP64, the
 *                   iterative-refinement correction of the mixed-precision sweep: every slice's
 *                   complex64 inverse is kept, and X = H^-1 E_1 (H the block-tridiagonal matrix of
 *                   src/waveguide.f90:424-440) is corrected `refine` times by the FP64 residual
 *                   E_1 - H X solved with those inverses; the FP64 eigensolver follows.  Passes hold
 *                   3 nx np^2 complex128 per realization (the pass size follows from the free device
 *                   memory); slower than a complex128 sweep at N = 128 (see rg_refine.hip);
 *   "refine_cap"    >= 0 (default 0): at most this many realizations per refined pass (0: as many as
 *                   60% of the free device memory holds); setting it frees the refinement's buffers;
 *   "lanes"         1 one stream, 2 the batch split in halves on two streams;
 *   "eigen"         transmission eigenvalues: 0 eigenvalues of t^H t (Householder tridiagonalisation
 *                   and Sturm bisection); 1 one-sided Jacobi SVD of t; 2 Householder bidiagonalisation
 *                   of t and Sturm bisection of its singular values (orders <= 256);
 *                   3 (default) 2 for complex128 sweeps, 0 for complex64 sweeps.  The SVD forms keep
 *                   each sigma within a few n eps sigma_max, as zgesvd (src/base_utils.f90:341-380);
 *                   t^H t loses closed channels (T ~ eps T_max absolute, src/base_utils.f90:385);
 *   "profile_mask"  bit k set: rg_plan_profile times the launches of entry k of
 *                   rg_plan_kernel_times (default: all bits);
 *   "graph"         1: seeded and explicit-potential sweeps run as captured HIP graphs (captured
 *                   on first use per batch size, precision, output pointer and options; replayed
 *                   with the seed read from device memory), when profiling is off and one lane is
 *                   active; 0 (default): direct launches;
 *   "adjoint_precision" RG_FP64 (default): the adjoint sweep in complex128; RG_MIXED_FP32: complex64
 *                   state and products (FP32 MFMA), loss, terminal adjoint and gradient in FP64.
 * RG_EINVAL for an unknown key or value.  Waits for the plan's stream first. */
rg_status rg_plan_set_option(rg_plan *plan, const char *key, int32_t value);
rg_status rg_plan_get_option(const rg_plan *plan, const char *key, int32_t *value);
/* Human-readable text of a status code (static storage). */
const char *rg_status_string(int32_t code);
/* Library version string. */
const char *rg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RECURGREEN_HIP_H */
