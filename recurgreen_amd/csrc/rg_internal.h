// rg_internal.h -- shared declarations of the gfx950 RGF library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/recurgreen_hip.h"

// ------------------------------------------------------------------------------------------------
// Complex double helpers (interleaved double2 = Fortran complex(real64)).
// ------------------------------------------------------------------------------------------------
__host__ __device__ inline double2 cmk(double x, double y) { return make_double2(x, y); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cneg(double2 a) { return cmk(-a.x, -a.y); }
__host__ __device__ inline double2 cconj(double2 a) { return cmk(a.x, -a.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b)
{
    return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc - a*b with fused multiply-adds
__device__ inline double2 cfnms(double2 acc, double2 a, double2 b)
{
    acc.x = fma(-a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(-a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
    return acc;
}
// acc + a*b
__device__ inline double2 cfma(double2 acc, double2 a, double2 b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
// 1/a with Smith's range reduction
__host__ __device__ inline double2 crecip(double2 a)
{
    if (fabs(a.x) >= fabs(a.y)) {
        double r = a.y / a.x, den = a.x + a.y * r;
        return cmk(1.0 / den, -r / den);
    }
    double r = a.x / a.y, den = a.x * r + a.y;
    return cmk(r / den, -1.0 / den);
}
__host__ __device__ inline double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }

// Element traits of the complex128 / complex64 batched products: the real type, the MFMA
// accumulator, the 16x16x4 MFMA (both take A[i][k] from lane i + 16k and B[k][j] from lane j + 16k)
// and where result register v of lane group g lands (f64: row 4v + g; f32: row 4g + v).
template <class E> struct rg_mt;
template <> struct rg_mt<double2> {
    typedef double R;
    typedef double V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int v, int g) { return 4 * v + g; }
    static __device__ __forceinline__ double2 make(double x, double y) { return make_double2(x, y); }
    static __device__ __forceinline__ double2 ld(double2 v) { return v; }
};
template <> struct rg_mt<float2> {
    typedef float R;
    typedef float V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int v, int g) { return 4 * g + v; }
    static __device__ __forceinline__ float2 make(double x, double y) { return make_float2((float)x, (float)y); }
    static __device__ __forceinline__ double2 ld(float2 v) { return make_double2(v.x, v.y); }
};

// ------------------------------------------------------------------------------------------------
// Plan.
// ------------------------------------------------------------------------------------------------
struct rg_timer {
    std::string name;
    double ms = 0.0;
    int64_t launches = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
};

enum rg_kernel_id { RGK_PROJECT = 0, RGK_INVERSE, RGK_GEMM, RGK_SVD, RGK_HIST, RGK_COPY, RGK_REDO, RGK_COUNT };

// Slice coefficients as the projection leaves them: `parts` partial sums over disjoint ranges of
// transverse cells, cq[(part * nb + b) * ntab + q], to be added in part order and scaled by 1/ny.
struct rg_cqv {
    const double *cq;  // null: no disorder on this slice
    int ntab, nb, parts;
    double scale;
};

__device__ inline void rg_load_cq(double *s_cq, const rg_cqv &v, int b, int tid, int nthreads)
{
    for (int q = tid; q < v.ntab; q += nthreads) {
        double part[8];
#pragma unroll
        for (int k = 0; k < 8; k++) part[k] = k < v.parts ? v.cq[((size_t)k * v.nb + b) * v.ntab + q] : 0.0;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) s += part[k];  // part order: the same sum on every path
        s_cq[q] = s * v.scale;
    }
}

// Tile t of the lower triangle of a matrix of wv x wv tiles, as (I, J) with I >= J: the diagonal tiles
// first, then the off-diagonal ones by distance from the diagonal (the symmetric slice inversions).
__device__ __forceinline__ void rg_sym_tile(int t, int wv, int &I, int &J)
{
    if (t < wv) {
        I = J = t;
        return;
    }
    t -= wv;
    int d = 1;
    while (t >= wv - d) {
        t -= wv - d;
        d++;
    }
    J = t;
    I = t + d;
}

enum rg_eigen_method { RG_EIG_GRAM = 0, RG_EIG_JACOBI = 1, RG_EIG_BIDIAG = 2, RG_EIG_AUTO = 3 };

// RG_INV_SYM: verified diagonal pivots on the lower triangle (complex symmetric slice matrices,
// Dirichlet guides); RG_INV_NOPIV elsewhere.
enum rg_inverse_kernel { RG_INV_REGISTER = 0, RG_INV_LOOKAHEAD = 1, RG_INV_BLOCKED = 2, RG_INV_NOPIV = 3, RG_INV_SYM = 4 };

struct rg_plan {
    rg_params p{};
    int nmu = 0, ny = 0, nx = 0, neigv = 0, imin = 0, imax = 0, jmin = 0, jmax = 0, nprop = 0;
    int nxdiso = 0, nxfree = 0;
    double kl = 0, kdx = 0, dos = 0, stdu = 0;
    int np = 0;        // padded matrix order (multiple of 16)
    int ntab = 0;      // projection coefficients per slice (doubles)
    int ntp = 0;       // ntab padded to a multiple of 16 (table row length)
    int nchunk = 0;    // transverse cells in chunks of 16 (table rows = 16 * nchunk, zero padded)
    int cq_parts = 1;  // partial sums of the projection (transverse-cell split), at most
    int cq_parts_used = 1;  // partial sums written by the last projection launch (<= cq_parts)
    int project_wide_min = 1;  // smallest batch projected by k_project_gen / k_project_wide (plan option)
    int device = 0;
    hipStream_t stream = nullptr;
    int batch_max = 0;
    int upot_cap = 0;  // realizations the explicit-potential staging buffer holds

    // host copies of the mode constants (np entries, padded)
    std::vector<double2> h_mu, h_b, h_abound, h_abulk, h_sqrtmu;

    // device constants
    double2 *d_b = nullptr, *d_adiag = nullptr, *d_sqrtmu = nullptr;
    double *d_sqrtremu = nullptr;  // sqrt(Re mu), reflection/transmission relations (np)  // d_adiag: 3*np (bound, bulk, last)
    double *d_tab = nullptr;                                           // (16 nchunk) x ntp, zero padded
    // batch buffers
    double2 *d_gss = nullptr, *d_gs1[2] = {nullptr, nullptr}, *d_tmat = nullptr;
    float2 *d_gss32 = nullptr, *d_gs1_32[2] = {nullptr, nullptr};  // FP32 state of the mixed sweeps
    // reflection sweep (G11, G1s ping-pong) for up to refl_cap realizations per pass
    int refl_cap = 0;
    double2 *d_g11 = nullptr, *d_g1s[2] = {nullptr, nullptr};
    // adjoint sweep workspace (rg_adjoint_reserve): checkpoints every adj_seg slices, one segment
    // of recomputed states, five adjoint matrices, the gradient staging, for adj_cap realizations
    int adj_cap = 0, adj_seg = 0, adj_nck = 0;
    int adj_prec = 0;  // RG_FP64: complex128 adjoint; RG_MIXED_FP32: complex64 state and products
    // FP64 refinement of the complex64 sweeps (plan option refine: passes after the first): for passes of
    // ref_cap realizations, every slice's g_x, X and R (complex128, np x np) and coefficients; ref_vcap V blocks
    int refine = 0, ref_cap = 0, ref_vcap = 0;
    int refine_cap = 0;  // plan option: at most this many realizations per refined pass (0: from the free memory)
    double2 *d_ref_g = nullptr, *d_ref_x = nullptr, *d_ref_r = nullptr, *d_ref_w = nullptr, *d_ref_v = nullptr;
    double *d_ref_cq = nullptr;
    double2 *d_ck = nullptr, *d_seg = nullptr, *d_adj = nullptr;
    double *d_loss = nullptr, *d_grad = nullptr;
    double2 *d_bconj = nullptr;  // conj(b) (np)
    double2 *d_adj_work = nullptr;  // T_max above order 128: Jacobi work X and V (2 adj_cap np^2)
    // tprofile workspace (rg_tprofile_reserve), tp_cap realizations per pass: Gss and Gs1 of every
    // slice (nx slabs of tp_cap np^2 each), the SVD work X and V, W = s U, two products Y, the
    // eigenvalues (tp_cap x nmu) and the intensities (tp_cap x nx x nmu)
    int tp_cap = 0;
    double2 *d_tp_g = nullptr, *d_tp_z = nullptr, *d_tp_x = nullptr, *d_tp_v = nullptr, *d_tp_w = nullptr;
    double2 *d_tp_y[2] = {nullptr, nullptr};
    double *d_tp_t = nullptr, *d_tp_i = nullptr;
    double2 *d_tp_psi = nullptr;  // the transverse fields of the wavefunction pass (nx x ny)
    double *d_cq = nullptr, *d_teig = nullptr, *d_upot = nullptr;
    double *d_us = nullptr;  // disorder of the current slice for the batch (batch_max x ny)
    int32_t *d_status = nullptr;
    int32_t *d_redo = nullptr;  // per realization: 1 = the verified-pivot inversion handed it back
    unsigned long long *d_hist = nullptr;  // counts + missing + dropped
    double *d_bins = nullptr;
    int hist_cap = 0;

    bool profiling = false;
    int profile_mask = ~0;  // bit k: record launches of timer k (rg_kernel_id) while profiling
    // slice-inversion kernel per state precision (RG_INV_*): FP64 state, complex64 state
    int inv_fp64 = 4, inv_mixed = 4;
    // transmission eigenvalues (option eigen, rg_eigen_method): RG_EIG_GRAM eig(t^H t), RG_EIG_JACOBI one-sided
    // Jacobi SVD of t, RG_EIG_BIDIAG Householder bidiagonalisation of t (order <= 256), RG_EIG_AUTO
    // the bidiagonal SVD for complex128 sweeps and the Gram eigensolver for complex64 sweeps
    int eigen = 3;
    // complex64 Dyson product (option dyson): 0 64 x 64 tiles, 1 one realization per workgroup on FP32
    // MFMA (NP >= 80), 2 the same on BF16 MFMA with three-way split operands (NP 96 and 128, else 1),
    // 3 one realization per workgroup with three real products per complex product (NP 96 and 128, else 1)
    int gemm_wg = 3;
    int ncu = 256;  // compute units of the device: the grid of the persistent kernels
    rg_timer timers[RGK_COUNT];
    // Captured sweeps (option "graph"): one executable graph per shape/options/output, replayed with
    // the seed read from the device word d_seed (seed_dev points at it while a sweep is captured).
    struct graph_entry {
        int nb, precision, opts;
        const double *upot;
        double *teig;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    // Uneven batches: cstride = np when every realization has its own mode constants (d_adiag
    // [r][3][np], d_b / d_sqrtmu [r][np]), 0 when they are shared; per-realization seeds and
    // disorder strengths for the noise (null: seed0 + r and stdu).
    // Slice inversion above np = 128 (rg_schur.hip): block work area of batch_max realizations,
    // views start schur_off realizations in; a zero diagonal and b = i for the block inversions.
    double2 *d_schur = nullptr, *d_schur_zero = nullptr, *d_schur_i = nullptr;
    int schur_off = 0;
    int cstride = 0;
    const int32_t *d_nx_r = nullptr;  // slice counts: slice cur_ix uses the last diagonal where nx_r == cur_ix
    int cur_ix = 0;
    const int64_t *d_seeds_r = nullptr;
    const double *d_stdu_r = nullptr;
    bool use_graph = false;
    std::vector<graph_entry> graphs;
    int64_t *d_seed = nullptr;
    const int64_t *seed_dev = nullptr;
    // Concurrent lanes: the batch split in halves, each swept on a stream of its own by a view of
    // this plan (same buffers, re-based pointers).  lanes = 1: one stream.
    int lanes = 1;
    // complex128 Dyson product: 1 one workgroup per realization at NP 80..128 (default), 2 the same
    // with three real products per complex product, 0 the 64 x 64-tiled kernel (option dyson_fp64)
    int dyson64 = 1;
    rg_plan *lane[2] = {nullptr, nullptr};
    hipEvent_t lane_ev[3] = {nullptr, nullptr, nullptr};
};

// Smallest half-batch worth a lane of its own.
constexpr int kLaneMin = 128;
// Most partial sums of the projection (rg_load_cq adds at most eight).
constexpr int kMaxCqParts = 8;

// rg_project.hip
void rg_launch_project(rg_plan *P, int nb, int ixd, int64_t seed0, const double *d_upot, int upot_nreal);
int rg_project_parts(int ny);
bool rg_project_supported(int ntab);
// rg_project_gen.hip: the seeded projection of one slice with the disorder generated in the kernel
// (3 to 17 column tiles); false (nothing launched) for other table widths
bool rg_launch_project_gen(rg_plan *P, int nb, int ixd, int64_t seed0);
// Kernel launchers (rg_kernels_f64.hip).
void rg_launch_inverse_f64(rg_plan *P, int nb, int slot, int first, bool with_disorder);
void rg_launch_gemm_f64(rg_plan *P, int nb, const double2 *gss, const double2 *gs1_old, double2 *gs1_new);
void rg_launch_fisher_lee_svd_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status,
                                  double2 *tmat_out);
void rg_launch_fisher_lee_svd_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status,
                                  double2 *tmat_out);
// rg_zgemm.hip: C = alpha diag(dl) op(A) diag(dk) op(B) diag(dr) + beta C (op: 0 = N, 1 = H), n = nmu;
// C must not alias A or B.
void rg_zgemm(rg_plan *P, int nb, double2 alpha, const double2 *A, int opA, const double2 *B, int opB, double2 beta,
              double2 *C, const double2 *dl, const double2 *dk, const double2 *dr);
// rg_adjoint.hip
void rg_cgemm32(rg_plan *P, int nb, double2 alpha, const float2 *A, int opA, const float2 *B, int opB, double2 beta,
                float2 *C, const double2 *dl, const double2 *dk, const double2 *dr);
// rg_adjoint.hip (complex128 and complex64 adjoint arrays)
void rg_launch_adj_terminal(rg_plan *P, int nb, const double2 *gs1, double *loss, double2 *sbar);
void rg_launch_adj_terminal(rg_plan *P, int nb, const float2 *gs1, double *loss, float2 *sbar);
void rg_launch_adj_terminal_tmax(rg_plan *P, int nb, const float2 *gs1, const double2 *xvec, int xstride, double *loss,
                                 float2 *sbar);
void rg_launch_adj_add(rg_plan *P, int nb, float2 *gbar, const float2 *sbar);
void rg_launch_adj_scale(rg_plan *P, int nb, float2 *gbar, const float2 *cbar);
void rg_launch_adj_grad(rg_plan *P, int nb, const float2 *cbar, int ixd, double *grad);
void rg_launch_adj_terminal_tmax(rg_plan *P, int nb, const double2 *gs1, const double2 *xvec, int xstride, double *loss,
                                 double2 *sbar);
void rg_launch_adj_add(rg_plan *P, int nb, double2 *gbar, const double2 *sbar);
void rg_launch_adj_scale(rg_plan *P, int nb, double2 *gbar, const double2 *cbar);
void rg_launch_adj_grad(rg_plan *P, int nb, const double2 *cbar, int ixd, double *grad);
// rg_refl.hip
void rg_launch_refl_trans_out(rg_plan *P, int nb, const double2 *g11, const double2 *gs1, double2 *rmat, double2 *tmat);
void rg_launch_deposition_out(rg_plan *P, int nb, const double2 *g, double2 *dep);
// rg_tprofile.hip
bool rg_tprofile_supported(int n);
void rg_launch_tprofile_svd(rg_plan *P, int nb, const double2 *gs1, int i0, int m, int j0, int nn, double2 *xwork,
                            double2 *vwork, double2 *w, double *tvals);
void rg_launch_colnorms(rg_plan *P, int nb, const double2 *y, int nv, double *out, size_t ostride);
void rg_launch_unit_column(rg_plan *P, double2 *w, int row);
// top eigenvector of the T_max Gram matrix for orders up to 256 (work X, V: np^2 complex128 per realization)
void rg_launch_topvec_jacobi(rg_plan *P, int nb, const double2 *gs1, double2 *xwork, double2 *vwork, double2 *xvec,
                             int xstride);
void rg_launch_topvec_jacobi(rg_plan *P, int nb, const float2 *gs1, double2 *xwork, double2 *vwork, double2 *xvec,
                             int xstride);
void rg_launch_modal_column(rg_plan *P, const double2 *y, double2 *psi);
void rg_launch_add_c128(rg_plan *P, double2 *a, const double2 *b, size_t count);
// rg_svd.hip: singular values of t by Householder bidiagonalisation, orders <= 256 (above 128 with t in
// the t buffer d_tmat)
bool rg_tsvd_supported(const rg_plan *P);
void rg_launch_tsvd_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status);
void rg_launch_tsvd_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status);
// Transmission eigenvalues of G(L;0) by the plan's method (rg_plan.hip)
void rg_launch_teig_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status);
void rg_launch_teig_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status);
// rg_eig.hip
void rg_launch_tdt_eig_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status);
void rg_launch_tdt_eig_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status);
bool rg_eig_supported(const rg_plan *P);
bool rg_topvec_supported(const rg_plan *P);  // top eigenvector (adjoint of T_max): order <= 128
// top eigenvector of t^H t (or t t^H when t has fewer rows), 128 complex slots per realization
void rg_launch_tdt_topvec_f64(rg_plan *P, int nb, const double2 *gs1, int32_t *status, double2 *xvec);
void rg_launch_tdt_topvec_f32(rg_plan *P, int nb, const float2 *gs1, int32_t *status, double2 *xvec);
// rg_gj.hip (only != nullptr: invert just the realizations with only[b] != 0)
void rg_launch_inverse_gj(rg_plan *P, int nb, int slot, int first, bool with_disorder, int precision,
                          const int32_t *only = nullptr);
// rg_nopiv.hip (complex64 state) and rg_nopiv64.hip (complex128 state): blocked Gauss-Jordan on verified
// diagonal pivots; realizations whose check fails are flagged in redo (inverted again by rg_launch_inverse_gj)
void rg_launch_inverse_nopiv(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo);
void rg_launch_inverse_nopiv64(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo);
void rg_launch_inverse_sym64(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo);
void rg_launch_inverse_sym32(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo);
// rg_blocked.hip
void rg_launch_inverse_blocked(rg_plan *P, int nb, int slot, int first, bool with_disorder, int precision);
void rg_launch_gemm_sp(rg_plan *P, int nb, const float2 *gss, const float2 *gs1_old, float2 *gs1_new, int precision);
// rg_schur.hip: slice inversion for 128 < np <= 256 (2 x 2 Schur complement on the smaller kernels);
// rg_schur_elems: complex128 elements of block work area per realization.
void rg_launch_inverse_schur(rg_plan *P, int nb, int slot, int first, bool with_disorder, int precision);
// C = alpha A B + beta C per realization, complex128 (rg_schur.hip)
void rg_zgemm(rg_plan *P, int nb, int m, int n, int k, double2 alpha, const double2 *A, int lda, size_t sA,
              const double2 *B, int ldb, size_t sB, double2 beta, double2 *C, int ldc, size_t sC);
// FP64 refinement of the complex64 sweeps (rg_refine.hip): buffers (sets ref_cap), the store of slice ix
// after its step, and the refinement passes, returning G(L;0) (nb x np x np complex128)
rg_status rg_refine_reserve(rg_plan *P);
void rg_refine_release(rg_plan *P);
void rg_refine_store(rg_plan *P, int nb, int ix, bool with_disorder, const float2 *gs1);
const double2 *rg_refine_solve(rg_plan *P, int nb);
size_t rg_schur_elems(int np);
// rg_dyson32.hip: the FP32 Dyson product with one realization per workgroup; false if NP is not
// one of 80, 96, 112, 128
bool rg_launch_gemm_wg(rg_plan *P, int nb, const float2 *gss, const float2 *gs1_old, float2 *gs1_new);
// the same with three real products per complex product, persistent (NP 96 and 128; false otherwise)
bool rg_launch_gemm_3m(rg_plan *P, int nb, const float2 *gss, const float2 *gs1_old, float2 *gs1_new);
// rg_dyson64.hip: the complex128 product, one workgroup per realization (NP 80..128; false otherwise)
bool rg_launch_gemm_wg64(rg_plan *P, int nb, const double2 *gss, const double2 *gs1_old, double2 *gs1_new);
bool rg_launch_gemm_bf3(rg_plan *P, int nb, const float2 *gss, const float2 *gs1_old, float2 *gs1_new);
void rg_launch_copy_f2(rg_plan *P, const float2 *src, float2 *dst, size_t n);
void rg_launch_histogram(rg_plan *P, const double *d_samples, int64_t n, const double *d_bins, int nbins,
                         unsigned long long *d_counts);
// the same on any stream (counts + missing + dropped added into d_counts, nbins <= 4096)
void rg_launch_histogram_on(hipStream_t stream, const double *d_samples, int64_t n, const double *d_bins, int nbins,
                            unsigned long long *d_counts);
void rg_launch_potential(rg_plan *P, int nb, int64_t seed0, double *d_upot);

rg_cqv rg_cq_view(const rg_plan *P, int nb, bool with_disorder);

// Timing helpers (rg_plan.hip).
void rg_timer_begin(rg_plan *P, int id, hipEvent_t *stop_out);
void rg_timer_end(rg_plan *P, int id, hipEvent_t stop);

#include "rg_cos.h"

// Host-side Philox4x32-10 (shared with the device generator: same constants and rounds).
__host__ __device__ inline void rg_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Standard normals of disorder cells (ix, 2p) and (ix, 2p + 1) of realization `seed`: two 53-bit uniforms
// from the Philox block of counter (p, ix, stream, 0), u1 in (0,1], u2 in [0,1), then both outputs of the
// Box-Muller transform, z0 = sqrt(-2 log u1) cos(2 pi u2) and z1 = sqrt(-2 log u1) sin(2 pi u2) (the
// reference keeps the cosine output only, src/random_normal.f90:64): one block, logarithm and square root
// per two cells.
__host__ __device__ inline void rg_normal_pair(uint64_t seed, uint32_t stream, int ix, int p, double &z0, double &z1)
{
    uint32_t c[4] = {(uint32_t)p, (uint32_t)ix, stream, 0u};
    rg_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32) ^ 0x52474E31u);
    uint64_t a = ((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6);
    uint64_t b = ((uint64_t)(c[2] >> 5) << 26) | (uint64_t)(c[3] >> 6);
    double u1 = (double)(a + 1) * 0x1.0p-53;
    double u2 = (double)b * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    double s, co;
    rg_sincos_0_2pi(2.0 * 3.14159265358979323846 * u2, &s, &co);
    z0 = r * co;
    z1 = r * s;
}

// Cell (ix, iy): the cosine output of pair iy / 2 for even iy, the sine output for odd iy.
__host__ __device__ inline double rg_normal_cell(uint64_t seed, uint32_t stream, int ix, int iy)
{
    double z0, z1;
    rg_normal_pair(seed, stream, ix, iy >> 1, z0, z1);
    return (iy & 1) ? z1 : z0;
}
