// rg_svd.hip -- transmission eigenvalues as squared singular values of t, orders <= 256 (gfx950).
//
// src/waveguide.f90:443-449 forms t_ij = 2i sqrt(mu_i) G_ij(L;0) sqrt(mu_j) over the aperture ranges
// and takes T = sigma(t)^2 from zgesvd (src/base_utils.f90:341-380): Householder bidiagonalisation,
// then the singular values of the real bidiagonal.  This kernel follows the same two steps in FP64,
// one workgroup per realization:
//   1. t (or t^H when t has more columns than rows) in registers: 512 threads, thread (r, c) =
//      (tid % 16, tid / 16) holds rows r + 16 i (i < 8) and columns c + 32 j (j < 4), cyclically,
//      so the 16 owners of a column are one 16-lane row group of a wavefront;
//   2. for k = 0, 1, ...: the left reflector of column k (the zlarfg vector, applied as H^H to the
//      columns right of k; the column products reduce inside a row group by lane shuffles), then
//      the right reflector of row k (applied from the right to the rows below k; the row products
//      reduce across the 32 column groups: lane shuffles inside a wavefront, LDS across them).  Only
//      |beta| of each reflector is kept: a bidiagonal's singular values do not depend on the phases
//      of its entries;
//   3. the singular values as the positive eigenvalues of the Golub-Kahan tridiagonal (zero diagonal,
//      off-diagonals d1 e1 d2 e2 ... dn) by Sturm-count multisection, four threads a value (five
//      sub-intervals an iteration), each value found by its index, so the output is sorted.
// The Householder steps are backward stable and Sturm counts of the Golub-Kahan matrix resolve the
// bidiagonal's singular values to high relative accuracy: every sigma is within a few n eps
// sigma_max of the exact singular value of t, as zgesvd's.  The eigenvalues of t^H t (rg_eig.hip)
// carry n eps sigma_max^2 instead, an error of eps sigma_max^2 / sigma^2 relative to a closed
// channel's T = sigma^2.
#include "rg_internal.h"

namespace {

constexpr int SV_NT = 512;      // threads per workgroup
constexpr int SV_MAXN = 128;    // largest order (rows and columns of t)
constexpr int SV_RI = 8;        // row slots per thread (16 row groups)
constexpr int SV_CJ = 4;        // column slots per thread (32 column groups)
constexpr int SV_NW = SV_NT / 64;
constexpr int SV_ITERS = 80;    // multisection steps: 5^-80 of the Gershgorin range
constexpr double SV_EPS = 2.220446049250313e-16;

__device__ inline double2 ld_sv(const double2 *p) { return *p; }
__device__ inline double2 ld_sv(const float2 *p)
{
    const float2 v = *p;
    return cmk((double)v.x, (double)v.y);
}

// Sum over the 16 lanes of a row group.
__device__ inline double red16(double v)
{
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 16);
    return v;
}

// Sum over lanes l, l^16, l^32, l^48 (the four column groups of a wavefront with the same r).
__device__ inline double red4(double v)
{
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// 1/q to full FP64 accuracy: hardware reciprocal and two Newton steps.
__device__ inline double rcp_sv(double q)
{
    double r = __builtin_amdgcn_rcp(q);
    r = fma(fma(-q, r, 1.0), r, r);
    r = fma(fma(-q, r, 1.0), r, r);
    return r;
}

// zlarfg for (alpha, x) with sum |x|^2 = ss: tau, the scale of x in v = (1, x scale), and |beta|.
__device__ inline void larfg(double2 al, double ss, double2 &tau, double2 &scale, double &babs)
{
    if (ss == 0.0 && al.y == 0.0) {  // H = I
        tau = cmk(0.0, 0.0);
        scale = cmk(0.0, 0.0);
        babs = fabs(al.x);
        return;
    }
    const double nrm = sqrt(fma(al.x, al.x, fma(al.y, al.y, ss)));
    const double beta = al.x >= 0.0 ? -nrm : nrm;
    tau = cmk((beta - al.x) / beta, -al.y / beta);
    scale = crecip(cmk(al.x - beta, al.y));
    babs = nrm;
}

// Eigenvalues below x of the Golub-Kahan tridiagonal (zero diagonal, squared off-diagonals f2[0..n-2]).
__device__ inline int gk_count(const double *f2, int n, double x, double pivmin)
{
    double q = -x;
    if (fabs(q) < pivmin) q = -pivmin;
    int cnt = q < 0.0;
    for (int i = 1; i < n; i++) {
        q = -x - f2[i - 1] * rcp_sv(q);
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

struct sv_lds {
    double2 v[SV_MAXN];          // left reflector, by row
    double2 wp[SV_NW][SV_MAXN];  // right reflector: row products, one partial per wavefront
    double2 w[SV_MAXN];
    double ssp[SV_NW];
    double2 alpha, tau;
    double f[2 * SV_MAXN];   // |d0| |e0| |d1| |e1| ...
    double f2[2 * SV_MAXN];  // their squares
    double hi;
    int bad;
};

// Step k = 16 KI + (k % 16) of the bidiagonalisation: row k sits in row slot KI of the threads r == k % 16,
// column k in column slot KI / 2 of column group k % 32, so every register index is a constant.
template <int KI>
__device__ __forceinline__ void sv_step(double2 (&a)[SV_RI][SV_CJ], sv_lds &L, int k, int R, int C, int r, int c, int lane,
                               int wave, int tid)
{
    constexpr int KJ = KI >> 1;
    const int kr = k & 15;
    const int src = (lane & ~15) | kr;
    // ---- left reflector: column k, rows k..R-1
    if (c == (k & 31)) {
        double ss = 0.0;
#pragma unroll
        for (int i = 0; i < SV_RI; i++) {
            const int row = r + 16 * i;
            if (row > k && row < R) ss = fma(a[i][KJ].x, a[i][KJ].x, fma(a[i][KJ].y, a[i][KJ].y, ss));
        }
        ss = red16(ss);
        double2 al;
        al.x = __shfl(a[KI][KJ].x, src, 64);  // row k on lane r == k % 16 of the group
        al.y = __shfl(a[KI][KJ].y, src, 64);
        double2 tau, scale;
        double babs;
        larfg(al, ss, tau, scale, babs);
#pragma unroll
        for (int i = 0; i < SV_RI; i++) {
            const int row = r + 16 * i;
            L.v[row] = row == k ? cmk(1.0, 0.0) : (row > k && row < R ? cmul(a[i][KJ], scale) : cmk(0.0, 0.0));
        }
        if (r == 0) {
            L.tau = tau;
            L.f[2 * k] = babs;
        }
    }
    __syncthreads();
    {
        const double2 tau = L.tau;
        if (tau.x != 0.0 || tau.y != 0.0) {  // uniform
            double2 v[SV_RI];
#pragma unroll
            for (int i = 0; i < SV_RI; i++) v[i] = L.v[r + 16 * i];
            const double2 ct = cconj(tau);
#pragma unroll
            for (int j = 0; j < SV_CJ; j++) {
                double2 acc = cmk(0.0, 0.0);
#pragma unroll
                for (int i = 0; i < SV_RI; i++) acc = cfma(acc, cconj(v[i]), a[i][j]);
                acc.x = red16(acc.x);
                acc.y = red16(acc.y);
                if (c + 32 * j > k) {
                    const double2 f = cmul(ct, acc);  // A -= conj(tau) v (v^H A)
#pragma unroll
                    for (int i = 0; i < SV_RI; i++) a[i][j] = cfnms(a[i][j], v[i], f);
                }
            }
        }
    }
    if (k + 1 >= C) return;
    // ---- right reflector: row k, columns k+1..C-1 (column k+1 in slot KJ, or KJ + 1 when k + 1 = 32 (KJ + 1))
    {
        double ss = 0.0;
        if (r == kr) {
#pragma unroll
            for (int j = 0; j < SV_CJ; j++) {
                const int col = c + 32 * j;
                if (col > k + 1 && col < C) ss = fma(a[KI][j].x, a[KI][j].x, fma(a[KI][j].y, a[KI][j].y, ss));
            }
            if (c == ((k + 1) & 31)) {
                constexpr int KJ1 = KJ + 1 < SV_CJ ? KJ + 1 : SV_CJ - 1;
                L.alpha = cconj(((k + 1) & 31) ? a[KI][KJ] : a[KI][KJ1]);
            }
        }
        ss = red4(ss);
        if (lane == kr) L.ssp[wave] = ss;
    }
    __syncthreads();
    double ss = 0.0;
#pragma unroll
    for (int q = 0; q < SV_NW; q++) ss += L.ssp[q];
    double2 tau, scale;
    double babs;
    larfg(L.alpha, ss, tau, scale, babs);
    if (tid == 0) L.f[2 * k + 1] = babs;
    if (tau.x == 0.0 && tau.y == 0.0) return;  // uniform; the next barrier orders the LDS reuse
    double2 v[SV_CJ];
#pragma unroll
    for (int j = 0; j < SV_CJ; j++) {
        double2 y;
        y.x = __shfl(a[KI][j].x, src, 64);  // row k of this column group
        y.y = __shfl(a[KI][j].y, src, 64);
        const int col = c + 32 * j;
        v[j] = col == k + 1 ? cmk(1.0, 0.0) : (col > k + 1 && col < C ? cmul(cconj(y), scale) : cmk(0.0, 0.0));
    }
#pragma unroll
    for (int i = 0; i < SV_RI; i++) {
        double2 acc = cmk(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < SV_CJ; j++) acc = cfma(acc, a[i][j], v[j]);
        acc.x = red4(acc.x);
        acc.y = red4(acc.y);
        if (lane < 16) L.wp[wave][r + 16 * i] = acc;
    }
    __syncthreads();
    if (tid < SV_MAXN) {
        double2 sw = L.wp[0][tid];
#pragma unroll
        for (int q = 1; q < SV_NW; q++) sw = cadd(sw, L.wp[q][tid]);
        L.w[tid] = sw;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SV_RI; i++) {
        if (r + 16 * i > k) {
            const double2 f = cmul(tau, L.w[r + 16 * i]);  // A -= tau (A v) v^H
#pragma unroll
            for (int j = 0; j < SV_CJ; j++)
                if (c + 32 * j > k) a[i][j] = cfnms(a[i][j], f, cconj(v[j]));
        }
    }
}

template <int KI>
__device__ __forceinline__ void sv_steps(double2 (&a)[SV_RI][SV_CJ], sv_lds &L, int R, int C, int r, int c, int lane, int wave,
                                int tid)
{
    if constexpr (KI < SV_RI) {
        for (int kr = 0; kr < 16; kr++) {
            const int k = 16 * KI + kr;
            if (k >= C) return;
            sv_step<KI>(a, L, k, R, C, r, c, lane, wave, tid);
        }
        sv_steps<KI + 1>(a, L, R, C, r, c, lane, wave, tid);
    }
}

template <class E>
__global__ __launch_bounds__(SV_NT, 1) void k_tsvd(const E *__restrict__ gs1, int NP, int imin, int imax, int jmin,
                                                   int jmax, const double2 *__restrict__ sqrtmu,
                                                   double *__restrict__ teig, int neigv, int32_t *__restrict__ status,
                                                   int cs)
{
    sqrtmu += (size_t)blockIdx.x * cs;  // cs = np: constants of each realization, else shared
    __shared__ sv_lds L;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = tid & 15, c = tid >> 4;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const bool tr = nn > m;  // work on t^H: rows >= columns
    const int R = tr ? nn : m, C = tr ? m : nn;
    const E *G = gs1 + (size_t)b * NP * NP;

    if (tid == 0) L.bad = 0;
    __syncthreads();
    double2 a[SV_RI][SV_CJ];
    bool bad = false;
#pragma unroll
    for (int i = 0; i < SV_RI; i++)
#pragma unroll
        for (int j = 0; j < SV_CJ; j++) {
            const int row = r + 16 * i, col = c + 32 * j;
            double2 v = cmk(0.0, 0.0);
            if (row < R && col < C) {
                const int gi = imin - 1 + (tr ? col : row), gj = jmin - 1 + (tr ? row : col);
                const double2 g = cmul(cmul(sqrtmu[gi], ld_sv(G + (size_t)gi * NP + gj)), sqrtmu[gj]);
                const double2 t = cmk(-2.0 * g.y, 2.0 * g.x);  // 2i g
                v = tr ? cconj(t) : t;
                bad |= !isfinite(v.x) || !isfinite(v.y);
            }
            a[i][j] = v;
        }
    if (bad) L.bad = 1;
    __syncthreads();
    if (L.bad) {
        if (tid == 0) status[b] = RG_ENONFINITE;
        return;
    }
    sv_steps<0>(a, L, R, C, r, c, lane, wave, tid);
    __syncthreads();

    // Golub-Kahan tridiagonal of order 2C.
    const int n2 = 2 * C;
    for (int i = tid; i < n2 - 1; i += SV_NT) L.f2[i] = L.f[i] * L.f[i];
    if (tid == 0) {
        double hi = 0.0;
        for (int i = 0; i < n2; i++) hi = fmax(hi, (i > 0 ? L.f[i - 1] : 0.0) + (i + 1 < n2 ? L.f[i] : 0.0));
        L.hi = hi;
    }
    __syncthreads();
    double f2max = 0.0;
    for (int i = 0; i < n2 - 1; i++) f2max = fmax(f2max, L.f2[i]);
    const double pivmin = 1e-290 * fmax(1.0, f2max);
    const int j = tid >> 2, h = tid & 3;  // thread quad j finds the (j+1)-th largest singular value
    if (j < C && j < neigv) {             // uniform inside a quad
        const int J = n2 - 1 - j;         // its ascending index in the Golub-Kahan spectrum (+-sigma)
        double lo = 0.0, up = L.hi * (1.0 + 8.0 * SV_EPS) + pivmin;
        const int q0 = lane & ~3;
        for (int it = 0; it < SV_ITERS; it++) {
            const double step = (up - lo) * 0.2;
            const int cnt = gk_count(L.f2, n2, lo + (h + 1) * step, pivmin);
            const int c0 = __shfl(cnt, q0, 64), c1 = __shfl(cnt, q0 + 1, 64);
            const int c2 = __shfl(cnt, q0 + 2, 64), c3 = __shfl(cnt, q0 + 3, 64);
            const double x0 = lo + step, x1 = lo + 2.0 * step, x2 = lo + 3.0 * step, x3 = lo + 4.0 * step;
            if (c0 > J) up = x0;
            else if (c1 > J) { lo = x0; up = x1; }
            else if (c2 > J) { lo = x1; up = x2; }
            else if (c3 > J) { lo = x2; up = x3; }
            else lo = x3;
            if (up - lo <= 2.0 * SV_EPS * fmax(fabs(lo), fabs(up)) + pivmin) break;
        }
        if (h == 0) {
            const double sv = 0.5 * (lo + up);
            const double tv = sv * sv;
            // complex64 state: T projected onto [0, 1], where every exact eigenvalue lies (unitarity,
            // no gain), as the Gram eigensolver does for those sweeps
            teig[(size_t)b * neigv + j] = std::is_same<E, float2>::value ? fmin(tv, 1.0) : tv;
        }
    }
}


// ------------------------------------------------------------------------------------------------
// Orders 129..256: the same reduction with t in a per-realization work area (column-major, leading
// dimension NP, streamed through the caches) instead of registers: 1024 threads, thread (r, c) =
// (lane % 32, tid / 32) owns rows r + 32 i and columns c + 32 j (i, j < 8), so the 32 owners of a
// column are one half of a wavefront (column sums by lane shuffles) and the 32 owners of a row are
// spread over the 16 wavefronts (row sums through LDS).  Each reflector reads the trailing elements
// twice (its products, then its update).
// ------------------------------------------------------------------------------------------------
constexpr int SG_NT = 1024;
constexpr int SG_MAXN = 256;
constexpr int SG_NW = SG_NT / 64;

struct sg_lds {
    double2 v[SG_MAXN];
    double2 wp[SG_NW][SG_MAXN];
    double2 w[SG_MAXN];
    double ssp[SG_NW];
    double2 alpha, tau;
    double f[2 * SG_MAXN];
    double f2[2 * SG_MAXN];
    double hi;
    int bad;
};

// Sum over the 32 lanes of a half wavefront.
__device__ inline double red32(double v)
{
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 32);
    return v;
}

template <class E>
__global__ __launch_bounds__(SG_NT) void k_tsvd_g(const E *__restrict__ gs1, int NP, int imin, int imax, int jmin,
                                                  int jmax, const double2 *__restrict__ sqrtmu,
                                                  double2 *__restrict__ work, double *__restrict__ teig, int neigv,
                                                  int32_t *__restrict__ status, int cs)
{
    sqrtmu += (size_t)blockIdx.x * cs;
    __shared__ sg_lds L;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 31, c = tid >> 5;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const bool tr = nn > m;
    const int R = tr ? nn : m, C = tr ? m : nn;
    const E *G = gs1 + (size_t)b * NP * NP;
    double2 *X = work + (size_t)b * NP * NP;  // X[col * NP + row]

    if (tid == 0) L.bad = 0;
    __syncthreads();
    bool bad = false;
    for (int j = 0; j < 8; j++) {
        const int col = c + 32 * j;
        for (int i = 0; i < 8; i++) {
            const int row = r + 32 * i;
            if (row >= R || col >= C) continue;
            const int gi = imin - 1 + (tr ? col : row), gj = jmin - 1 + (tr ? row : col);
            const double2 g = cmul(cmul(sqrtmu[gi], ld_sv(G + (size_t)gi * NP + gj)), sqrtmu[gj]);
            const double2 t = cmk(-2.0 * g.y, 2.0 * g.x);  // 2i g
            const double2 v = tr ? cconj(t) : t;
            bad |= !isfinite(v.x) || !isfinite(v.y);
            X[(size_t)col * NP + row] = v;
        }
    }
    if (bad) L.bad = 1;
    __syncthreads();
    if (L.bad) {
        if (tid == 0) status[b] = RG_ENONFINITE;
        return;
    }

    for (int k = 0; k < C; k++) {
        const int kr = k & 31;
        // ---- left reflector: column k, rows k..R-1 (the 32 threads of column group k % 32)
        if (c == (k & 31)) {
            const double2 *xk = X + (size_t)k * NP;
            double ss = 0.0;
            double2 xv[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int row = r + 32 * i;
                xv[i] = row < R ? xk[row] : cmk(0.0, 0.0);
                if (row > k && row < R) ss = fma(xv[i].x, xv[i].x, fma(xv[i].y, xv[i].y, ss));
            }
            ss = red32(ss);
            const double2 own = xk[k];  // every lane of the group reads the diagonal entry
            double2 tau, scale;
            double babs;
            larfg(own, ss, tau, scale, babs);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int row = r + 32 * i;
                L.v[row] = row == k ? cmk(1.0, 0.0) : (row > k && row < R ? cmul(xv[i], scale) : cmk(0.0, 0.0));
            }
            if (r == 0) {
                L.tau = tau;
                L.f[2 * k] = babs;
            }
        }
        __syncthreads();
        {
            const double2 tau = L.tau;
            if (tau.x != 0.0 || tau.y != 0.0) {  // uniform
                double2 v[8];
#pragma unroll
                for (int i = 0; i < 8; i++) v[i] = L.v[r + 32 * i];
                const double2 ct = cconj(tau);
                for (int j = 0; j < 8; j++) {
                    const int col = c + 32 * j;  // uniform in a half wavefront
                    double2 acc = cmk(0.0, 0.0);
                    double2 *xc = X + (size_t)col * NP;
                    if (col > k && col < C)
#pragma unroll
                        for (int i = 0; i < 8; i++) {
                            const int row = r + 32 * i;
                            if (row >= k && row < R) acc = cfma(acc, cconj(v[i]), xc[row]);
                        }
                    acc.x = red32(acc.x);
                    acc.y = red32(acc.y);
                    if (col > k && col < C) {
                        const double2 f = cmul(ct, acc);  // A -= conj(tau) v (v^H A)
#pragma unroll
                        for (int i = 0; i < 8; i++) {
                            const int row = r + 32 * i;
                            if (row >= k && row < R) xc[row] = cfnms(xc[row], v[i], f);
                        }
                    }
                }
            }
        }
        __syncthreads();  // the left update of row k is visible before its right reflector reads it
        if (k + 1 >= C) break;
        // ---- right reflector: row k, columns k+1..C-1
        double2 rk[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int col = c + 32 * j;
            rk[j] = col > k && col < C ? X[(size_t)col * NP + k] : cmk(0.0, 0.0);
        }
        {
            double ss = 0.0;
            if (r == kr) {
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (c + 32 * j > k + 1) ss = fma(rk[j].x, rk[j].x, fma(rk[j].y, rk[j].y, ss));
                if (c == ((k + 1) & 31)) L.alpha = cconj(X[(size_t)(k + 1) * NP + k]);
            }
            ss += __shfl_xor(ss, 32, 64);
            if (lane == kr) L.ssp[wave] = ss;
        }
        __syncthreads();
        double ss = 0.0;
        for (int q = 0; q < SG_NW; q++) ss += L.ssp[q];
        double2 tau, scale;
        double babs;
        larfg(L.alpha, ss, tau, scale, babs);
        if (tid == 0) L.f[2 * k + 1] = babs;
        if (tau.x == 0.0 && tau.y == 0.0) continue;  // uniform; the next barrier orders the LDS reuse
        double2 v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int col = c + 32 * j;
            v[j] = col == k + 1 ? cmk(1.0, 0.0) : (col > k + 1 && col < C ? cmul(cconj(rk[j]), scale) : cmk(0.0, 0.0));
        }
        for (int i = 0; i < 8; i++) {
            const int row = r + 32 * i;
            double2 acc = cmk(0.0, 0.0);
            if (row > k && row < R)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int col = c + 32 * j;
                    if (col > k && col < C) acc = cfma(acc, X[(size_t)col * NP + row], v[j]);
                }
            acc.x += __shfl_xor(acc.x, 32, 64);
            acc.y += __shfl_xor(acc.y, 32, 64);
            if (lane < 32) L.wp[wave][row] = acc;
        }
        __syncthreads();
        if (tid < SG_MAXN) {
            double2 sw = L.wp[0][tid];
            for (int q = 1; q < SG_NW; q++) sw = cadd(sw, L.wp[q][tid]);
            L.w[tid] = sw;
        }
        __syncthreads();
        for (int i = 0; i < 8; i++) {
            const int row = r + 32 * i;
            if (row > k && row < R) {
                const double2 f = cmul(tau, L.w[row]);  // A -= tau (A v) v^H
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int col = c + 32 * j;
                    if (col > k && col < C) {
                        double2 *p = X + (size_t)col * NP + row;
                        *p = cfnms(*p, f, cconj(v[j]));
                    }
                }
            }
        }
        __syncthreads();  // the right update is complete before the next left reflector reads column k+1
    }
    __syncthreads();

    const int n2 = 2 * C;
    for (int i = tid; i < n2 - 1; i += SG_NT) L.f2[i] = L.f[i] * L.f[i];
    if (tid == 0) {
        double hi = 0.0;
        for (int i = 0; i < n2; i++) hi = fmax(hi, (i > 0 ? L.f[i - 1] : 0.0) + (i + 1 < n2 ? L.f[i] : 0.0));
        L.hi = hi;
    }
    __syncthreads();
    double f2max = 0.0;
    for (int i = 0; i < n2 - 1; i++) f2max = fmax(f2max, L.f2[i]);
    const double pivmin = 1e-290 * fmax(1.0, f2max);
    const int j = tid >> 2, h = tid & 3;
    if (j < C && j < neigv) {
        const int J = n2 - 1 - j;
        double lo = 0.0, up = L.hi * (1.0 + 8.0 * SV_EPS) + pivmin;
        const int q0 = lane & ~3;
        for (int it = 0; it < SV_ITERS; it++) {
            const double step = (up - lo) * 0.2;
            const int cnt = gk_count(L.f2, n2, lo + (h + 1) * step, pivmin);
            const int c0 = __shfl(cnt, q0, 64), c1 = __shfl(cnt, q0 + 1, 64);
            const int c2 = __shfl(cnt, q0 + 2, 64), c3 = __shfl(cnt, q0 + 3, 64);
            const double x0 = lo + step, x1 = lo + 2.0 * step, x2 = lo + 3.0 * step, x3 = lo + 4.0 * step;
            if (c0 > J) up = x0;
            else if (c1 > J) { lo = x0; up = x1; }
            else if (c2 > J) { lo = x1; up = x2; }
            else if (c3 > J) { lo = x2; up = x3; }
            else lo = x3;
            if (up - lo <= 2.0 * SV_EPS * fmax(fabs(lo), fabs(up)) + pivmin) break;
        }
        if (h == 0) {
            const double sv = 0.5 * (lo + up);
            const double tv = sv * sv;
            teig[(size_t)b * neigv + j] = std::is_same<E, float2>::value ? fmin(tv, 1.0) : tv;
        }
    }
}

}  // namespace

bool rg_tsvd_supported(const rg_plan *P)
{
    const int m = P->imax - P->imin + 1, nn = P->jmax - P->jmin + 1;
    return (m > nn ? m : nn) <= SG_MAXN;
}

void rg_launch_tsvd_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    const int m = P->imax - P->imin + 1, nn = P->jmax - P->jmin + 1;
    if ((m > nn ? m : nn) <= SV_MAXN)
        hipLaunchKernelGGL(k_tsvd<double2>, dim3(nb), dim3(SV_NT), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, teig, P->neigv, status, P->cstride);
    else  // t in the t buffer (batch_max x np^2 complex128), free while eigenvalues are computed
        hipLaunchKernelGGL(k_tsvd_g<double2>, dim3(nb), dim3(SG_NT), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, P->d_tmat, teig, P->neigv, status, P->cstride);
    rg_timer_end(P, RGK_SVD, stop);
}

void rg_launch_tsvd_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    const int m = P->imax - P->imin + 1, nn = P->jmax - P->jmin + 1;
    if ((m > nn ? m : nn) <= SV_MAXN)
        hipLaunchKernelGGL(k_tsvd<float2>, dim3(nb), dim3(SV_NT), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, teig, P->neigv, status, P->cstride);
    else
        hipLaunchKernelGGL(k_tsvd_g<float2>, dim3(nb), dim3(SG_NT), 0, P->stream, gs1, P->np, P->imin, P->imax,
                           P->jmin, P->jmax, P->d_sqrtmu, P->d_tmat, teig, P->neigv, status, P->cstride);
    rg_timer_end(P, RGK_SVD, stop);
}
