// rg_kernels_f64.hip -- complex128 kernels of the recursive Green function sweep on gfx950.
//
// One realization of the disorder is one row of the batch.  Per slice ix the sweep of
// src/waveguide.f90:431-440 runs as three launches over the whole batch:
//   k_project        disorder slice -> transform coefficients C(q)       (:366-380)
//   k_slice_inverse  Gss <- inv(A_ix - V_ix - B Gss B)                   (:434-435)
//   k_dyson_gemm     Gs1 <- -Gss B Gs1                                   (:438)
// and once per realization
//   k_fisher_lee_svd t = 2i sqrt(mu) Gs1 sqrt(mu), T = sigma(t)^2        (:443-449)
// Matrices live in HBM as row-major np x np complex128 (np = matrix order padded to 16).
#include "rg_internal.h"

namespace {

// ------------------------------------------------------------------------------------------------
// Disorder slice projection.  C(q) = (1/ny) sum_i u_i tab[i][q] for the ntab table columns:
// Dirichlet tab[i][q] = cos(q pi y_i), q = 0..2N  (V_mn = C(|m-n|) - C(m+n));
// periodic  tab[i][2q] = cos(2 pi q y_i), tab[i][2q+1] = sin(2 pi q y_i), q = 0..N-1
// (V_mn = D(n-m), D(q) = C(2q) + i C(2q+1)).  The slice u_i is either generated (Philox,
// realization seed seed0 + b) or read from an explicit potential batch upot(ix, iy, b).
// Work items (q, s) split the i range in KSPLIT fixed parts; partials are summed in a fixed
// order, so the result is deterministic.
// ------------------------------------------------------------------------------------------------
constexpr int KSPLIT = 4;

__global__ __launch_bounds__(256) void k_project(int ixd, int64_t seed0, const double *__restrict__ upot,
                                                 int nxdiso, int ny, double stdu,
                                                 const double *__restrict__ tab, int ntab,
                                                 double *__restrict__ cq)
{
    extern __shared__ double smem[];
    double *s_u = smem;            // ny
    double *s_part = smem + ny;    // KSPLIT * ntab
    const int b = blockIdx.x;
    for (int iy = threadIdx.x; iy < ny; iy += blockDim.x) {
        double u;
        if (upot) u = upot[(size_t)b * nxdiso * ny + (size_t)iy * nxdiso + ixd];
        else u = stdu * rg_normal_cell((uint64_t)(seed0 + b), 0u, ixd, iy);
        s_u[iy] = u;
    }
    __syncthreads();
    const int chunk = (ny + KSPLIT - 1) / KSPLIT;
    for (int w = threadIdx.x; w < ntab * KSPLIT; w += blockDim.x) {
        const int q = w % ntab, s = w / ntab;
        const int i0 = s * chunk, i1 = min(ny, i0 + chunk);
        double acc = 0.0;
        for (int i = i0; i < i1; i++) acc = fma(s_u[i], tab[(size_t)i * ntab + q], acc);
        s_part[s * ntab + q] = acc;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < ntab; q += blockDim.x) {
        double acc = 0.0;
        for (int s = 0; s < KSPLIT; s++) acc += s_part[s * ntab + q];
        cq[(size_t)b * ntab + q] = acc / ny;
    }
}

__global__ __launch_bounds__(256) void k_potential(int64_t seed0, int nxdiso, int ny, double stdu,
                                                   double *__restrict__ upot)
{
    const int b = blockIdx.y;
    const int64_t ncell = (int64_t)nxdiso * ny;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        int ix = (int)(c % nxdiso), iy = (int)(c / nxdiso);
        upot[(size_t)b * ncell + c] = stdu * rg_normal_cell((uint64_t)(seed0 + b), 0u, ix, iy);
    }
}

// Projected potential V_mn (0-based m, n < nmu) from the slice coefficients.
__device__ inline double2 slice_v(const double *s_cq, int boundtype, int m, int n)
{
    if (boundtype == RG_DIRICHLET) return cmk(s_cq[abs(m - n)] - s_cq[m + n + 2], 0.0);
    if (n >= m) return cmk(s_cq[2 * (n - m)], s_cq[2 * (n - m) + 1]);
    return cmk(s_cq[2 * (m - n)], -s_cq[2 * (m - n) + 1]);
}

// ------------------------------------------------------------------------------------------------
// Slice inversion by Gauss-Jordan elimination with partial pivoting, register resident.
//
// The np x np matrix (np = 16*TR) is spread over 4*np threads: thread (r, c) holds rows
// [r*TR, r*TR+TR) x columns [4c, 4c+4), r = tid % 16, c = tid / 16, so the 16 owners of a column
// sit in one wavefront and the pivot search is a 16-lane shuffle reduction.  Pivots are chosen
// like LAPACK's zgetrf (max |Re|+|Im| over the rows not yet used, first index on ties), rows are
// never swapped; the permutation is undone when the result is written: with p_k the pivot row of
// column k, inv(A)[i][p_l] = X[p_i][l].  Padded rows/columns (>= n) hold the identity and are
// skipped.  Two barriers per elimination step; the broadcast buffers alternate between steps.
// ------------------------------------------------------------------------------------------------
// One Gauss-Jordan elimination step on column k (column jj of the owning tile).
template <int TR, int TC>
__device__ inline void gj_step(double2 (&a)[TR][TC], const int k, const int jj, const int c, const int r,
                               const int row0, const int col0, unsigned &pivoted, double2 *s_row, double2 *s_col,
                               double2 &s_pv, int *s_perm, int &s_sing)
{
    if (c == k / TC) {  // the 16 owners of column k: one wavefront, 16 consecutive lanes
        double best = -1.0;
        int bidx = 0x7fffffff;
        double2 bval = cmk(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < TR; i++) {
            if (!((pivoted >> i) & 1u)) {
                const double v = cabs1(a[i][jj]);
                if (v > best) { best = v; bidx = row0 + i; bval = a[i][jj]; }
            }
        }
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) {
            const double ob = __shfl_xor(best, off, 16);
            const int oi = __shfl_xor(bidx, off, 16);
            const double ox = __shfl_xor(bval.x, off, 16);
            const double oy = __shfl_xor(bval.y, off, 16);
            if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; bval = cmk(ox, oy); }
        }
        if (r == 0) {
            s_perm[k] = bidx;
            s_pv = bval;
            if (!(best > 0.0)) s_sing = 1;
        }
    }
    __syncthreads();
    const int p = s_perm[k];
    double2 pv = s_pv;
    if (pv.x == 0.0 && pv.y == 0.0) pv = cmk(1.0, 0.0);
    const double2 d = crecip(pv);
    if (p / TR == r) {
        const int ip = p % TR;
        double2 rowv[TC];
#pragma unroll
        for (int j = 0; j < TC; j++) rowv[j] = a[0][j];
#pragma unroll
        for (int i = 1; i < TR; i++)
#pragma unroll
            for (int j = 0; j < TC; j++) {
                const bool sel = i == ip;
                rowv[j].x = sel ? a[i][j].x : rowv[j].x;
                rowv[j].y = sel ? a[i][j].y : rowv[j].y;
            }
#pragma unroll
        for (int j = 0; j < TC; j++) s_row[col0 + j] = (col0 + j == k) ? d : cmul(rowv[j], d);
    }
    if (c == k / TC) {
#pragma unroll
        for (int i = 0; i < TR; i++) s_col[row0 + i] = a[i][jj];
    }
    __syncthreads();
    double2 rv[TC];
#pragma unroll
    for (int j = 0; j < TC; j++) rv[j] = s_row[col0 + j];
    const bool own_col = c == k / TC;
#pragma unroll
    for (int i = 0; i < TR; i++) {
        const int gi = row0 + i;
        const double2 m = s_col[gi];
        const bool is_p = gi == p;
#pragma unroll
        for (int j = 0; j < TC; j++) {
            double2 v = cfnms(a[i][j], m, rv[j]);
            if (own_col && j == jj) v = cneg(cmul(m, d));
            a[i][j].x = is_p ? rv[j].x : v.x;
            a[i][j].y = is_p ? rv[j].y : v.y;
        }
    }
    if (p / TR == r) pivoted |= 1u << (p % TR);
}

template <int TR>
__global__ __launch_bounds__(64 * TR) void k_slice_inverse(double2 *__restrict__ gss, const double *__restrict__ cq,
                                                             int ntab, int boundtype,
                                                             const double2 *__restrict__ adiag,
                                                             const double2 *__restrict__ bvec, int n, int first,
                                                             int32_t *__restrict__ status)
{
    constexpr int NP = 16 * TR, TC = 4, NT = 64 * TR;
    // staging height: a power of two with STAGE_ROWS * NP <= 2048 (32 KiB); NP is a multiple of 16
    constexpr int STAGE_ROWS = NP <= 16 ? 16 : (NP <= 32 ? 64 : (NP <= 64 ? 32 : 16));
    __shared__ double s_cq[2 * 256 + 8];
    __shared__ double2 s_row[2][NP];
    __shared__ double2 s_col[2][NP];
    __shared__ double2 s_pv[2];
    __shared__ int s_perm[NP];
    __shared__ int s_iperm[NP];
    __shared__ int s_sing;
    __shared__ double2 s_stage[STAGE_ROWS * NP];

    const int b = blockIdx.x, tid = threadIdx.x;
    const int r = tid & 15, c = tid >> 4;
    const int row0 = r * TR, col0 = c * TC;
    double2 *G = gss + (size_t)b * NP * NP;

    if (cq)
        for (int q = tid; q < ntab; q += NT) s_cq[q] = cq[(size_t)b * ntab + q];
    if (tid == 0) s_sing = 0;
    __syncthreads();

    // M = (diag(a) - V) - B Gss B on the n x n block, identity on the padding.
    double2 a[TR][TC];
#pragma unroll
    for (int i = 0; i < TR; i++) {
        const int gi = row0 + i;
#pragma unroll
        for (int j = 0; j < TC; j++) {
            const int gj = col0 + j;
            double2 m;
            if (gi < n && gj < n) {
                m = gi == gj ? adiag[gi] : cmk(0.0, 0.0);
                if (cq) m = csub(m, slice_v(s_cq, boundtype, gi, gj));
                if (!first) m = csub(m, cmul(cmul(bvec[gi], G[(size_t)gi * NP + gj]), bvec[gj]));
            } else {
                m = cmk(gi == gj ? 1.0 : 0.0, 0.0);
            }
            a[i][j] = m;
        }
    }

    unsigned pivoted = 0u;
    // Elimination steps, unrolled by the tile width so that the column inside a tile (jj) is a
    // compile-time index of the register array.
    for (int kb = 0; kb < n; kb += TC) {
#pragma unroll
        for (int jj = 0; jj < TC; jj++) {
            const int k = kb + jj;
            if (k >= n) break;
            gj_step<TR, TC>(a, k, jj, c, r, row0, col0, pivoted, s_row[k & 1], s_col[k & 1], s_pv[k & 1], s_perm, s_sing);
        }
    }
    for (int k = n + tid; k < NP; k += NT) s_perm[k] = k;
    __syncthreads();
    for (int k = tid; k < NP; k += NT) s_iperm[s_perm[k]] = k;
    if (tid == 0 && s_sing) status[b] = RG_ESINGULAR;

    // Write inv(A): physical row r' -> logical row iperm[r'], column j <- physical column iperm[j].
    for (int ch = 0; ch < NP; ch += STAGE_ROWS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TR; i++) {
            const int gi = row0 + i;
            if (gi >= ch && gi < ch + STAGE_ROWS)
#pragma unroll
                for (int j = 0; j < TC; j++) s_stage[(gi - ch) * NP + col0 + j] = a[i][j];
        }
        __syncthreads();
        for (int e = tid; e < STAGE_ROWS * NP; e += NT) {
            const int rr = e / NP, j = e % NP;
            if (ch + rr >= NP) break;
            const int dst = s_iperm[ch + rr];
            G[(size_t)dst * NP + j] = s_stage[rr * NP + s_iperm[j]];
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Dyson update Gs1_new = -Gss diag(b) Gs1_old, batched, LDS-tiled 64x64 output per workgroup,
// 4x4 complex outputs per thread, k-chunks of 16 (src/waveguide.f90:438).
// ------------------------------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(256) void k_dyson_gemm(const double2 *__restrict__ gss, const double2 *__restrict__ gs1,
                                                     double2 *__restrict__ out, const double2 *__restrict__ bvec, int n)
{
    constexpr int TT = (NP + 63) / 64;
    __shared__ double2 As[16][64 + 1];
    __shared__ double2 Bs[16][64 + 1];
    const int b = blockIdx.y, ti = blockIdx.x / TT, tj = blockIdx.x % TT;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const double2 *A = gss + (size_t)b * NP * NP;
    const double2 *B = gs1 + (size_t)b * NP * NP;
    double2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = cmk(0.0, 0.0);

    for (int k0 = 0; k0 < n; k0 += 16) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = threadIdx.x + 256 * e;
            {
                const int rr = idx >> 4, kk = idx & 15;
                const int gr = ti * 64 + rr, gk = k0 + kk;
                As[kk][rr] = (gr < NP && gk < n) ? A[(size_t)gr * NP + gk] : cmk(0.0, 0.0);
            }
            {
                const int kk = idx >> 6, cc = idx & 63;
                const int gk = k0 + kk, gc = tj * 64 + cc;
                Bs[kk][cc] = (gk < n && gc < NP) ? cmul(bvec[gk], B[(size_t)gk * NP + gc]) : cmk(0.0, 0.0);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            double2 av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = cfma(acc[i][j], av[i], bv[j]);
        }
        __syncthreads();
    }
    double2 *C = out + (size_t)b * NP * NP;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int gr = ti * 64 + ty + 16 * i;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int gc = tj * 64 + tx + 16 * j;
            if (gr < NP && gc < NP) C[(size_t)gr * NP + gc] = cneg(acc[i][j]);
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Fisher-Lee relation and singular values (src/waveguide.f90:443-449, src/base_utils.f90:341-380).
// t_ac = 2i sqrt(mu_i) G_ij sqrt(mu_j) (complex sqrt, :443) over the filtered ranges, then the
// one-sided (Hestenes) Jacobi method on the columns of t (or of t^H when t is wide), in the
// round-robin parallel ordering: every round rotates cols/2 disjoint column pairs, one pair per
// 16-lane group.  T_j = ||col_j||^2, sorted in decreasing order.  The column-major work matrix
// lives in HBM/L2 (ld = np) in the tmat workspace.
// ------------------------------------------------------------------------------------------------
__device__ inline void rr_pair(int round, int pr, int P, int &p, int &q)
{
    if (pr == 0) { p = round; q = P; }
    else { p = (round + pr) % P; q = (round - pr + P) % P; }
    if (p > q) { int t = p; p = q; q = t; }
}

__global__ __launch_bounds__(256) void k_fisher_lee_svd(const double2 *__restrict__ gs1, int NP, int imin, int imax,
                                                         int jmin, int jmax, const double2 *__restrict__ sqrtmu,
                                                         double2 *__restrict__ work, double *__restrict__ teig,
                                                         int neigv, int32_t *__restrict__ status, int max_sweeps)
{
    __shared__ int s_rot;
    __shared__ double s_t[512];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const bool tr = nn > m;
    const int rows = tr ? nn : m, cols = tr ? m : nn;
    const int cols_e = cols + (cols & 1);
    const double2 *G = gs1 + (size_t)b * NP * NP;
    double2 *X = work + (size_t)b * NP * NP;

    for (int e = tid; e < m * nn; e += blockDim.x) {
        const int ar = e / nn, ac = e % nn;
        const int i = imin - 1 + ar, j = jmin - 1 + ac;
        const double2 v = cmul(cmul(sqrtmu[i], G[(size_t)i * NP + j]), sqrtmu[j]);
        const double2 t = cmk(-2.0 * v.y, 2.0 * v.x);  // 2i * v
        if (!tr) X[(size_t)ac * NP + ar] = t;
        else X[(size_t)ar * NP + ac] = cconj(t);
    }
    if (cols_e != cols)
        for (int i = tid; i < rows; i += blockDim.x) X[(size_t)cols * NP + i] = cmk(0.0, 0.0);
    __syncthreads();

    const double tol = 2.220446049250313e-16 * sqrt((double)rows);
    const int grp = tid >> 4, lane = tid & 15, ngrp = blockDim.x >> 4;
    const int P = cols_e - 1;
    int sweep = 0;
    for (; sweep < max_sweeps; sweep++) {
        if (tid == 0) s_rot = 0;
        __syncthreads();
        for (int round = 0; round < P; round++) {
            for (int pr = grp; pr < cols_e / 2; pr += ngrp) {
                int p, q;
                rr_pair(round, pr, P, p, q);
                double2 *ap = X + (size_t)p * NP, *aq = X + (size_t)q * NP;
                double al = 0.0, be = 0.0, gx = 0.0, gy = 0.0;
                for (int i = lane; i < rows; i += 16) {
                    const double2 x = ap[i], y = aq[i];
                    al = fma(x.x, x.x, fma(x.y, x.y, al));
                    be = fma(y.x, y.x, fma(y.y, y.y, be));
                    gx = fma(x.x, y.x, fma(x.y, y.y, gx));   // conj(x) * y
                    gy = fma(x.x, y.y, fma(-x.y, y.x, gy));
                }
#pragma unroll
                for (int off = 8; off >= 1; off >>= 1) {
                    al += __shfl_xor(al, off, 16);
                    be += __shfl_xor(be, off, 16);
                    gx += __shfl_xor(gx, off, 16);
                    gy += __shfl_xor(gy, off, 16);
                }
                const double g = sqrt(gx * gx + gy * gy);
                if (g == 0.0 || g <= tol * sqrt(al * be)) continue;
                if (lane == 0) s_rot = 1;
                const double zeta = (be - al) / (2.0 * g);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                const double2 e = cmk(gx / g, -gy / g);  // exp(-i arg gamma)
                for (int i = lane; i < rows; i += 16) {
                    const double2 x = ap[i], y = cmul(e, aq[i]);
                    ap[i] = cmk(cs * x.x - sn * y.x, cs * x.y - sn * y.y);
                    aq[i] = cmk(sn * x.x + cs * y.x, sn * x.y + cs * y.y);
                }
            }
            __syncthreads();
        }
        if (!s_rot) break;
        __syncthreads();
    }
    for (int j = grp; j < cols; j += ngrp) {
        double s2 = 0.0;
        const double2 *aj = X + (size_t)j * NP;
        for (int i = lane; i < rows; i += 16) s2 = fma(aj[i].x, aj[i].x, fma(aj[i].y, aj[i].y, s2));
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) s2 += __shfl_xor(s2, off, 16);
        if (lane == 0) s_t[j] = s2;
    }
    __syncthreads();
    bool bad = false;
    for (int j = tid; j < cols; j += blockDim.x) {
        const double v = s_t[j];
        int rank = 0;
        for (int k2 = 0; k2 < cols; k2++) {
            const double w = s_t[k2];
            rank += (w > v) || (w == v && k2 < j);
        }
        if (rank < neigv) teig[(size_t)b * neigv + rank] = v;
        if (!(v == v) || v > 1e300) bad = true;
    }
    if (bad) status[b] = RG_ENONFINITE;
    if (tid == 0 && sweep == max_sweeps && s_rot && status[b] == RG_OK) status[b] = RG_ENOCONV;
}

// t written in the ABI layout: column-major m x nn per realization.
__global__ __launch_bounds__(256) void k_fisher_lee_out(const double2 *__restrict__ gs1, int NP, int imin, int imax,
                                                        int jmin, int jmax, const double2 *__restrict__ sqrtmu,
                                                        double2 *__restrict__ tout)
{
    const int b = blockIdx.x;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const double2 *G = gs1 + (size_t)b * NP * NP;
    for (int e = threadIdx.x; e < m * nn; e += blockDim.x) {
        const int ar = e % m, ac = e / m;
        const int i = imin - 1 + ar, j = jmin - 1 + ac;
        const double2 v = cmul(cmul(sqrtmu[i], G[(size_t)i * NP + j]), sqrtmu[j]);
        tout[(size_t)b * m * nn + e] = cmk(-2.0 * v.y, 2.0 * v.x);
    }
}

// ------------------------------------------------------------------------------------------------
// Histogram with the counting rules of src/base_utils.f90:491-516; bins by binary search
// (first j with x < bins[j+1], identical to the reference's linear scan on sorted edges).
// out[0..nbins-1] counts, out[nbins] missing, out[nbins+1] NaN dropped.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_histogram(const double *__restrict__ s, int64_t n,
                                                   const double *__restrict__ bins, int nbins,
                                                   unsigned long long *__restrict__ out)
{
    extern __shared__ double hsm[];
    double *s_bins = hsm;
    unsigned *s_cnt = reinterpret_cast<unsigned *>(hsm + nbins + 1);
    for (int i = threadIdx.x; i <= nbins; i += blockDim.x) s_bins[i] = bins[i];
    for (int i = threadIdx.x; i < nbins + 2; i += blockDim.x) s_cnt[i] = 0u;
    __syncthreads();
    const double lo = s_bins[0], hi = s_bins[nbins];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = s[i];
        int slot;
        if (x != x) slot = nbins + 1;
        else if (x < lo || hi < x) slot = nbins;
        else if (x == hi) slot = nbins - 1;
        else {
            int a = 0, z = nbins - 1;
            while (a < z) {
                const int mid = (a + z) >> 1;
                if (x < s_bins[mid + 1]) z = mid;
                else a = mid + 1;
            }
            slot = a;
        }
        atomicAdd(&s_cnt[slot], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins + 2; i += blockDim.x)
        if (s_cnt[i]) atomicAdd(&out[i], (unsigned long long)s_cnt[i]);
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Launchers.
// ------------------------------------------------------------------------------------------------
void rg_launch_project(rg_plan *P, int nb, int ixd, int64_t seed0, const double *d_upot, int upot_nreal)
{
    (void)upot_nreal;
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_PROJECT, &stop);
    const size_t shm = sizeof(double) * ((size_t)P->ny + (size_t)KSPLIT * P->ntab);
    hipLaunchKernelGGL(k_project, dim3(nb), dim3(256), shm, P->stream, ixd, seed0, d_upot, P->nxdiso, P->ny,
                       P->stdu, P->d_tab, P->ntab, P->d_cq);
    rg_timer_end(P, RGK_PROJECT, stop);
}

void rg_launch_potential(rg_plan *P, int nb, int64_t seed0, double *d_upot)
{
    const int64_t ncell = (int64_t)P->nxdiso * P->ny;
    int gx = (int)((ncell + 255) / 256);
    if (gx > 1024) gx = 1024;
    hipLaunchKernelGGL(k_potential, dim3(gx, nb), dim3(256), 0, P->stream, seed0, P->nxdiso, P->ny, P->stdu, d_upot);
}

void rg_launch_inverse_f64(rg_plan *P, int nb, int slot, int first, bool with_disorder)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_INVERSE, &stop);
    const double *cq = with_disorder ? P->d_cq : nullptr;
    const double2 *ad = P->d_adiag + (size_t)slot * P->np;
    const int TR = P->np / 16;
#define RG_INV_CASE(T)                                                                                    \
    case T:                                                                                               \
        hipLaunchKernelGGL(k_slice_inverse<T>, dim3(nb), dim3(64 * T), 0, P->stream, P->d_gss, cq, P->ntab, \
                           P->p.boundtype, ad, P->d_b, P->nmu, first, P->d_status);                     \
        break;
    switch (TR) {
        RG_INV_CASE(1)
        RG_INV_CASE(2)
        RG_INV_CASE(3)
        RG_INV_CASE(4)
        RG_INV_CASE(5)
        RG_INV_CASE(6)
        RG_INV_CASE(7)
        RG_INV_CASE(8)
        default: break;
    }
#undef RG_INV_CASE
    rg_timer_end(P, RGK_INVERSE, stop);
}

void rg_launch_gemm_f64(rg_plan *P, int nb, const double2 *gss, const double2 *gs1_old, double2 *gs1_new)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_GEMM, &stop);
    const int np = P->np;
    const int tt = (np + 63) / 64;
#define RG_GEMM_CASE(T)                                                                                     \
    case 16 * T:                                                                                            \
        hipLaunchKernelGGL(k_dyson_gemm<16 * T>, dim3(tt * tt, nb), dim3(256), 0, P->stream, gss, gs1_old, gs1_new, \
                           P->d_b, P->nmu);                                                                 \
        break;
    switch (np) {
        RG_GEMM_CASE(1)
        RG_GEMM_CASE(2)
        RG_GEMM_CASE(3)
        RG_GEMM_CASE(4)
        RG_GEMM_CASE(5)
        RG_GEMM_CASE(6)
        RG_GEMM_CASE(7)
        RG_GEMM_CASE(8)
        default: break;
    }
#undef RG_GEMM_CASE
    rg_timer_end(P, RGK_GEMM, stop);
}

void rg_launch_fisher_lee_svd_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status,
                                  double2 *tmat_out)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    if (tmat_out)
        hipLaunchKernelGGL(k_fisher_lee_out, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, tmat_out);
    else
        hipLaunchKernelGGL(k_fisher_lee_svd, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, P->d_tmat, teig, P->neigv, status, 60);
    rg_timer_end(P, RGK_SVD, stop);
}

void rg_launch_histogram(rg_plan *P, const double *d_samples, int64_t n, const double *d_bins, int nbins,
                         unsigned long long *d_counts)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_HIST, &stop);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    const size_t shm = sizeof(double) * (nbins + 1) + sizeof(unsigned) * (nbins + 2);
    hipLaunchKernelGGL(k_histogram, dim3((unsigned)blocks), dim3(256), shm, P->stream, d_samples, n, d_bins, nbins,
                       d_counts);
    rg_timer_end(P, RGK_HIST, stop);
}
