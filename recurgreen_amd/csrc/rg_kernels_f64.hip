// rg_kernels_f64.hip -- complex128 kernels of the recursive Green function sweep on gfx950.
//
// One realization of the disorder is one row of the batch.  Per slice ix the sweep of
// src/waveguide.f90:431-440 runs as three launches over the whole batch:
//   k_slice_inverse  Gss <- inv(A_ix - V_ix - B Gss B)                   (:434-435)
//   k_dyson_gemm     Gs1 <- -Gss B Gs1                                   (:438)
// and once per realization
//   k_fisher_lee_svd t = 2i sqrt(mu) Gs1 sqrt(mu), T = sigma(t)^2        (:443-449)
// Matrices live in HBM as row-major np x np complex128 (np = matrix order padded to 16).
#include "rg_internal.h"

namespace {

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
// The np x np matrix (np = 16*TR) is spread over 4*np threads: thread (r, c), r = tid % 16,
// c = tid / 16, holds rows {r + 16 i, i < TR} x columns [4c, 4c+4), so the 16 owners of a column
// are 16 consecutive lanes of one wavefront (one DPP row) and the pivot search is a 4-stage DPP
// butterfly.  Pivots are chosen like LAPACK's zgetrf: max |Re|+|Im| over the rows not used yet,
// lowest row index on ties.  Rows are never swapped; with p_k the pivot row of column k the result
// is written as inv(A)[i][p_l] = X[p_i][l].
//
// Each step is one uniform rank-1 update a <- a - m rv^T with no per-element special cases:
// column k is first replaced by e_p, the scaled pivot row is rv = a[p,:]/pv (so rv_k = 1/pv), and
// the pivot row's multiplier is pv - 1, which turns a[p,:] into a[p,:]/pv.  Two barriers per step;
// the broadcast buffers alternate between steps.  Padded rows/columns (>= n) hold the identity and
// their steps are skipped.
// ------------------------------------------------------------------------------------------------
template <int CTRL>
__device__ inline int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// One butterfly stage of the (max value, min index) reduction over a 16-lane DPP row.
template <int CTRL>
__device__ inline void argmax_stage(double &best, int &bidx)
{
    const long long bits = __double_as_longlong(best);
    const int lo = dpp_i<CTRL>((int)(bits & 0xffffffffll));
    const int hi = dpp_i<CTRL>((int)(bits >> 32));
    const int oi = dpp_i<CTRL>(bidx);
    const double ob = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
}

// dst[0..TC) <- a[ip][0..TC) for a uniform ip, as a chain of scalar branches (no indexed register
// access, which would move the array to scratch).
template <int I, int TR, int TC>
__device__ inline void store_row(const double2 (&a)[TR][TC], const int ip, double2 *dst)
{
    if constexpr (I < TR) {
        if (ip == I) {
#pragma unroll
            for (int j = 0; j < TC; j++) dst[j] = a[I][j];
        } else {
            store_row<I + 1, TR, TC>(a, ip, dst);
        }
    }
}

template <int TR, int TC>
__device__ inline void gj_step(double2 (&a)[TR][TC], const int k, const int jj, const int c, const int r,
                               unsigned &pivoted, double2 *s_col, double2 *s_rt, double2 &s_pd, int *s_perm,
                               int &s_sing, const int np)
{
    if (c == k / TC) {  // the 16 owners of column k
        double best = -1.0;
        int bidx = 0x7fffffff;
#pragma unroll
        for (int i = 0; i < TR; i++) {
            if (!((pivoted >> i) & 1u)) {
                const double v = cabs1(a[i][jj]);
                if (v > best) { best = v; bidx = r + 16 * i; }
            }
        }
        argmax_stage<0xB1>(best, bidx);   // quad_perm [1,0,3,2]
        argmax_stage<0x4E>(best, bidx);   // quad_perm [2,3,0,1]
        argmax_stage<0x141>(best, bidx);  // row_half_mirror
        argmax_stage<0x140>(best, bidx);  // row_mirror
#pragma unroll
        for (int i = 0; i < TR; i++) {
            const int gi = r + 16 * i;
            const double2 x = a[i][jj];
            const bool isp = gi == bidx;
            s_col[gi] = isp ? cmk(x.x - 1.0, x.y) : x;
            if (isp) {
                s_perm[k] = gi;
                const bool sing = !(best > 0.0);
                s_pd = sing ? cmk(1.0, 0.0) : crecip(x);
                if (sing) s_sing = 1;
            }
            a[i][jj] = cmk(isp ? 1.0 : 0.0, 0.0);
        }
    }
    __syncthreads();
    const int p = s_perm[k];
    const double2 d = s_pd;
    const int ip = p >> 4;  // wave-uniform: the pivot row is register row ip of row group p % 16
    if (r == (p & 15)) store_row<0, TR, TC>(a, ip, s_rt + TC * c);
    __syncthreads();
    double2 rv[TC];
#pragma unroll
    for (int j = 0; j < TC; j++) rv[j] = cmul(s_rt[TC * c + j], d);
#pragma unroll
    for (int i = 0; i < TR; i++) {
        const double2 m = s_col[r + 16 * i];
#pragma unroll
        for (int j = 0; j < TC; j++) a[i][j] = cfnms(a[i][j], m, rv[j]);
    }
    if (r == (p & 15)) pivoted |= 1u << ip;
}

template <int TR>
__global__ __launch_bounds__(64 * TR) void k_slice_inverse(double2 *__restrict__ gss, const rg_cqv cqv,
                                                             int boundtype,
                                                             const double2 *__restrict__ adiag,
                                                             const double2 *__restrict__ bvec, int n, int first,
                                                             int32_t *__restrict__ status, int cs,
                                                             const int32_t *__restrict__ nxr, int ix)
{
    constexpr int NP = 16 * TR, TC = 4, NT = 64 * TR;
    adiag += (size_t)blockIdx.x * 3 * cs;  // cs = np: constants of each realization, else shared
    if (nxr && nxr[blockIdx.x] == ix) adiag += cs;  // this realization's last slice: its last diagonal
    bvec += (size_t)blockIdx.x * cs;
    // staging height for the write-out: a power of two with STAGE_ROWS * NP <= 2048 (32 KiB)
    constexpr int STAGE_ROWS = NP <= 16 ? 16 : (NP <= 32 ? 64 : (NP <= 64 ? 32 : 16));
    constexpr int RT = 2 * NP;  // two pivot-row buffers
    constexpr int SHARED = RT > STAGE_ROWS * NP ? RT : STAGE_ROWS * NP;
    __shared__ double s_cq[2 * 256 + 8];
    __shared__ double2 s_buf[SHARED];  // pivot row groups during elimination, staging afterwards
    __shared__ double2 s_col[2][NP];
    __shared__ double2 s_pd[2];
    __shared__ int s_perm[NP];
    __shared__ int s_iperm[NP];
    __shared__ int s_sing;

    const int b = blockIdx.x, tid = threadIdx.x;
    const int r = tid & 15, c = tid >> 4;
    const int col0 = c * TC;
    double2 *G = gss + (size_t)b * NP * NP;

    if (cqv.cq) rg_load_cq(s_cq, cqv, b, tid, NT);
    if (tid == 0) s_sing = 0;
    __syncthreads();

    // M = (diag(a) - V) - B Gss B on the n x n block, identity on the padding.
    double2 a[TR][TC];
#pragma unroll
    for (int i = 0; i < TR; i++) {
        const int gi = r + 16 * i;
#pragma unroll
        for (int j = 0; j < TC; j++) {
            const int gj = col0 + j;
            double2 m;
            if (gi < n && gj < n) {
                m = gi == gj ? adiag[gi] : cmk(0.0, 0.0);
                if (cqv.cq) m = csub(m, slice_v(s_cq, boundtype, gi, gj));
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
            const int buf = k & 1;
            gj_step<TR, TC>(a, k, jj, c, r, pivoted, s_col[buf], s_buf + buf * NP, s_pd[buf], s_perm, s_sing,
                            NP);
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
            const int gi = r + 16 * i;
            if (gi >= ch && gi < ch + STAGE_ROWS)
#pragma unroll
                for (int j = 0; j < TC; j++) s_buf[(gi - ch) * NP + col0 + j] = a[i][j];
        }
        __syncthreads();
        for (int e = tid; e < STAGE_ROWS * NP; e += NT) {
            const int rr = e / NP, j = e % NP;
            if (ch + rr >= NP) break;
            const int dst = s_iperm[ch + rr];
            G[(size_t)dst * NP + j] = s_buf[rr * NP + s_iperm[j]];
        }
    }
}
// ------------------------------------------------------------------------------------------------
// Dyson update Gs1_new = -Gss diag(b) Gs1_old, batched, LDS-tiled 64x64 output per workgroup,
// 4x4 complex outputs per thread, k-chunks of 16 (src/waveguide.f90:438).
// ------------------------------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(256) void k_dyson_gemm(const double2 *__restrict__ gss, const double2 *__restrict__ gs1,
                                                     double2 *__restrict__ out, const double2 *__restrict__ bvec, int n,
                                                     int cs)
{
    constexpr int TT = (NP + 63) / 64;
    __shared__ double2 As[16][64 + 1];
    __shared__ double2 Bs[16][64 + 1];
    const int b = blockIdx.y, ti = blockIdx.x / TT, tj = blockIdx.x % TT;
    bvec += (size_t)b * cs;
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
// The same Dyson update on the FP64 matrix cores (v_mfma_f64_16x16x4_f64).  Lane layout of one
// 16x16x4 product (measured on gfx950): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D register v of lane l holds D[4v + l/16][l%16].  A complex product is four real ones:
// C_re += A_re B_re - A_im B_im, C_im += A_re B_im + A_im B_re.  A workgroup of 4 wavefronts
// computes a 64x64 output tile, each wavefront a 32x32 block (2x2 MFMA tiles, re and im
// accumulators); k advances in chunks of 16 staged through LDS as separate re/im planes, with
// diag(b) applied while staging Gs1.
// ------------------------------------------------------------------------------------------------
typedef double rg_d4 __attribute__((ext_vector_type(4)));

template <int NP>
__global__ __launch_bounds__(256) void k_dyson_gemm_mfma(const double2 *__restrict__ gss, const double2 *__restrict__ gs1,
                                                          double2 *__restrict__ out, const double2 *__restrict__ bvec, int n,
                                                          int cs)
{
    constexpr int TT = (NP + 63) / 64, LD = 64 + 2;  // padded LDS rows (bank spread)
    __shared__ double AsR[16][LD], AsI[16][LD];      // A^T chunk: [k][row]
    __shared__ double BsR[16][LD], BsI[16][LD];      // (b B) chunk: [k][col]
    const int b = blockIdx.y, ti0 = (blockIdx.x / TT) * 64, tj0 = (blockIdx.x % TT) * 64;
    bvec += (size_t)b * cs;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const double2 *A = gss + (size_t)b * NP * NP;
    const double2 *B = gs1 + (size_t)b * NP * NP;

    rg_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) { cr[x][y] = (rg_d4){0, 0, 0, 0}; ci[x][y] = (rg_d4){0, 0, 0, 0}; }

    for (int k0 = 0; k0 < n; k0 += 16) {
        {   // A chunk: thread -> row tid/4, k (tid%4)*4 .. +3
            const int rr = tid >> 2, kq = (tid & 3) * 4, gr = ti0 + rr;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gk = k0 + kq + e;
                const double2 v = (gr < NP && gk < n) ? A[(size_t)gr * NP + gk] : cmk(0.0, 0.0);
                AsR[kq + e][rr] = v.x;
                AsI[kq + e][rr] = v.y;
            }
        }
        {   // B chunk: thread -> k tid/16, cols (tid%16)*4 .. +3, scaled by b[k]
            const int kk = tid >> 4, cq = (tid & 15) * 4, gk = k0 + kk;
            const double2 bk = gk < n ? bvec[gk] : cmk(0.0, 0.0);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gc = tj0 + cq + e;
                const double2 v = (gk < n && gc < NP) ? cmul(bk, B[(size_t)gk * NP + gc]) : cmk(0.0, 0.0);
                BsR[kk][cq + e] = v.x;
                BsI[kk][cq + e] = v.y;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kr = ks * 4 + lk;
            double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                ar[x] = AsR[kr][wr + 16 * x + li];
                ai[x] = AsI[kr][wr + 16 * x + li];
                nai[x] = -ai[x];
                br[x] = BsR[kr][wc + 16 * x + li];
                bi[x] = BsI[kr][wc + 16 * x + li];
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    cr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], cr[x][y], 0, 0, 0);
                    cr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[x], bi[y], cr[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], ci[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], ci[x][y], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    double2 *C = out + (size_t)b * NP * NP;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int gr = ti0 + wr + 16 * x + 4 * v + lk, gc = tj0 + wc + 16 * y + li;
                if (gr < NP && gc < NP) C[(size_t)gr * NP + gc] = cmk(-cr[x][y][v], -ci[x][y][v]);
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

__device__ inline double2 ldc(const double2 *p) { return *p; }
__device__ inline double2 ldc(const float2 *p) { const float2 v = *p; return cmk((double)v.x, (double)v.y); }

template <class E>
__global__ __launch_bounds__(256) void k_fisher_lee_svd(const E *__restrict__ gs1, int NP, int imin, int imax,
                                                         int jmin, int jmax, const double2 *__restrict__ sqrtmu,
                                                         double2 *__restrict__ work, double *__restrict__ teig,
                                                         int neigv, int32_t *__restrict__ status, int max_sweeps, int cs)
{
    __shared__ int s_rot;
    __shared__ double s_t[512];
    const int b = blockIdx.x, tid = threadIdx.x;
    sqrtmu += (size_t)b * cs;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const bool tr = nn > m;
    const int rows = tr ? nn : m, cols = tr ? m : nn;
    const int cols_e = cols + (cols & 1);
    const E *G = gs1 + (size_t)b * NP * NP;
    double2 *X = work + (size_t)b * NP * NP;

    for (int e = tid; e < m * nn; e += blockDim.x) {
        const int ar = e / nn, ac = e % nn;
        const int i = imin - 1 + ar, j = jmin - 1 + ac;
        const double2 v = cmul(cmul(sqrtmu[i], ldc(G + (size_t)i * NP + j)), sqrtmu[j]);
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
        // complex64 state: T is projected onto [0, 1], where every exact eigenvalue lies
        // (unitarity, no gain), so FP32 round-off does not push open channels out of range
        if (rank < neigv) teig[(size_t)b * neigv + rank] = std::is_same<E, float2>::value ? fmin(v, 1.0) : v;
        if (!(v == v) || v > 1e300) bad = true;
    }
    if (bad) status[b] = RG_ENONFINITE;
    if (tid == 0 && sweep == max_sweeps && s_rot && status[b] == RG_OK) status[b] = RG_ENOCONV;
}

// t written in the ABI layout: column-major m x nn per realization.
template <class E>
__global__ __launch_bounds__(256) void k_fisher_lee_out(const E *__restrict__ gs1, int NP, int imin, int imax,
                                                        int jmin, int jmax, const double2 *__restrict__ sqrtmu,
                                                        double2 *__restrict__ tout, int cs)
{
    const int b = blockIdx.x;
    sqrtmu += (size_t)b * cs;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const E *G = gs1 + (size_t)b * NP * NP;
    for (int e = threadIdx.x; e < m * nn; e += blockDim.x) {
        const int ar = e % m, ac = e / m;
        const int i = imin - 1 + ar, j = jmin - 1 + ac;
        const double2 v = cmul(cmul(sqrtmu[i], ldc(G + (size_t)i * NP + j)), sqrtmu[j]);
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
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    const double2 *ad = P->d_adiag + (size_t)slot * P->np;
    const int TR = P->np / 16;
#define RG_INV_CASE(T)                                                                                    \
    case T:                                                                                               \
        hipLaunchKernelGGL(k_slice_inverse<T>, dim3(nb), dim3(64 * T), 0, P->stream, P->d_gss, cq, \
                           P->p.boundtype, ad, P->d_b, P->nmu, first, P->d_status, P->cstride, P->d_nx_r, P->cur_ix); \
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
    if (P->dyson64 && rg_launch_gemm_wg64(P, nb, gss, gs1_old, gs1_new)) {
        rg_timer_end(P, RGK_GEMM, stop);
        return;
    }
    const int np = P->np;
    const int tt = (np + 63) / 64;
#define RG_GEMM_CASE(T)                                                                                     \
    case 16 * T:                                                                                            \
        hipLaunchKernelGGL(k_dyson_gemm_mfma<16 * T>, dim3(tt * tt, nb), dim3(256), 0, P->stream, gss, gs1_old, gs1_new, \
                           P->d_b, P->nmu, P->cstride);                                                     \
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
        RG_GEMM_CASE(9)
        RG_GEMM_CASE(10)
        RG_GEMM_CASE(11)
        RG_GEMM_CASE(12)
        RG_GEMM_CASE(13)
        RG_GEMM_CASE(14)
        RG_GEMM_CASE(15)
        RG_GEMM_CASE(16)
        default: break;
    }
#undef RG_GEMM_CASE
    rg_timer_end(P, RGK_GEMM, stop);
}

template <class E>
static void launch_fisher_lee_svd(rg_plan *P, int nb, const E *gs1, double *teig, int32_t *status, double2 *tmat_out)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    if (tmat_out)
        hipLaunchKernelGGL(k_fisher_lee_out<E>, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, tmat_out, P->cstride);
    else
        hipLaunchKernelGGL(k_fisher_lee_svd<E>, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, P->d_tmat, teig, P->neigv, status, 60, P->cstride);
    rg_timer_end(P, RGK_SVD, stop);
}

void rg_launch_fisher_lee_svd_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status,
                                  double2 *tmat_out)
{
    launch_fisher_lee_svd<double2>(P, nb, gs1, teig, status, tmat_out);
}

void rg_launch_fisher_lee_svd_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status,
                                  double2 *tmat_out)
{
    launch_fisher_lee_svd<float2>(P, nb, gs1, teig, status, tmat_out);
}

void rg_launch_histogram_on(hipStream_t stream, const double *d_samples, int64_t n, const double *d_bins, int nbins,
                            unsigned long long *d_counts)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    const size_t shm = sizeof(double) * (nbins + 1) + sizeof(unsigned) * (nbins + 2);
    hipLaunchKernelGGL(k_histogram, dim3((unsigned)blocks), dim3(256), shm, stream, d_samples, n, d_bins, nbins,
                       d_counts);
}

void rg_launch_histogram(rg_plan *P, const double *d_samples, int64_t n, const double *d_bins, int nbins,
                         unsigned long long *d_counts)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_HIST, &stop);
    rg_launch_histogram_on(P->stream, d_samples, n, d_bins, nbins, d_counts);
    rg_timer_end(P, RGK_HIST, stop);
}
