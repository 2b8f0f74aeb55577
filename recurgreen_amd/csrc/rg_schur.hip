// rg_schur.hip -- slice inversion for 128 < np <= 256 by a 2x2 Schur complement (gfx950).
//
// The slice matrix M = A_x - V_x - B Gss B (src/waveguide.f90:431-436) of order np > 128 does not fit
// the register-resident inversions (a 256 x 256 complex64 matrix is the whole register file of a
// CU).  It is split at n1 = 128 into blocks [[A, B], [C, D]] (n2 = np - 128) and inverted as
//   A^-1;  X = A^-1 B;  S = D - C X;  Y = C A^-1;  S^-1;
//   M^-1 = [[A^-1 + X S^-1 Y, -X S^-1], [-S^-1 Y, S^-1]],
// two inversions of order <= 128 on the kernels of the smaller orders (pivoting inside A and inside
// S) and six products on k_cgemm, the same 8 np^3 flops as a direct inversion.  The blocks live in
// a per-realization work area; the result is assembled into Gss.  Both precisions: complex128
// (FP64 MFMA) and complex64 (FP32 MFMA).
//
// Complex symmetric slice matrices (Dirichlet guides, complex128 with the symmetric inversion kernel):
// C = B^T, so Y = C A^-1 = X^T and S^-1 Y = (X S^-1)^T, and A and S are symmetric.  Then
//   A^-1 and S^-1 on the symmetric kernel (lower triangles), X = A^-1 B, S = D - C X, Z = -X S^-1,
//   M^-1 = [[A^-1 - Z X^T, Z], [Z^T, S^-1]]:
// four products instead of six.  The top-left block is written symmetrized, (T + T^T)/2, and the
// bottom-left as the exact transpose of the top-right, so that Gss stays exactly symmetric (an
// asymmetric rounding component of Gss is not damped by the recursion on absorbing guides).
#include <type_traits>

#include "rg_internal.h"

namespace {

// C = alpha A B + beta C for one m x n output per realization (A m x k, B k x n; leading dimensions
// lda/ldb/ldc, realization strides sA/sB/sC in elements).  64 x 64 output tile per workgroup,
// 32 x 32 per wavefront, k in chunks of 16 staged through LDS as re/im planes.  beta == 0 does not
// read C.
// TB: B is read transposed, B(kk, j) = B[j ldb + kk].
template <class E, bool TB>
__global__ __launch_bounds__(256) void k_cgemm(int m, int n, int k, double2 alpha, const E *__restrict__ A, int lda,
                                               size_t sA, const E *__restrict__ B, int ldb, size_t sB, double2 beta,
                                               E *__restrict__ C, int ldc, size_t sC)
{
    using R = typename rg_mt<E>::R;
    using V4 = typename rg_mt<E>::V4;
    __shared__ R AsR[16][66], AsI[16][66], BsR[16][66], BsI[16][66];
    const int b = blockIdx.z, ti0 = blockIdx.y * 64, tj0 = blockIdx.x * 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, li = lane & 15, lg = lane >> 4;
    const E *Ab = A + (size_t)b * sA;
    const E *Bb = B + (size_t)b * sB;
    V4 cr[2][2], ci[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            cr[x][y] = V4{0, 0, 0, 0};
            ci[x][y] = V4{0, 0, 0, 0};
        }
    for (int k0 = 0; k0 < k; k0 += 16) {
        {
            const int rr = tid >> 2, kq = (tid & 3) * 4, gr = ti0 + rr;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gk = k0 + kq + e;
                E v{0, 0};
                if (gr < m && gk < k) v = Ab[(size_t)gr * lda + gk];
                AsR[kq + e][rr] = v.x;
                AsI[kq + e][rr] = v.y;
            }
        }
        if constexpr (TB) {  // 16 consecutive k of one column per 16 threads
            const int kk = tid & 15, cq = (tid >> 4) * 4, gk = k0 + kk;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gc = tj0 + cq + e;
                E v{0, 0};
                if (gk < k && gc < n) v = Bb[(size_t)gc * ldb + gk];
                BsR[kk][cq + e] = v.x;
                BsI[kk][cq + e] = v.y;
            }
        } else {
            const int kk = tid >> 4, cq = (tid & 15) * 4, gk = k0 + kk;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gc = tj0 + cq + e;
                E v{0, 0};
                if (gk < k && gc < n) v = Bb[(size_t)gk * ldb + gc];
                BsR[kk][cq + e] = v.x;
                BsI[kk][cq + e] = v.y;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kr = 4 * ks + lg;
            R ar[2], ai[2], nai[2], br[2], bi[2];
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
                    cr[x][y] = rg_mt<E>::mfma(ar[x], br[y], cr[x][y]);
                    cr[x][y] = rg_mt<E>::mfma(nai[x], bi[y], cr[x][y]);
                    ci[x][y] = rg_mt<E>::mfma(ar[x], bi[y], ci[x][y]);
                    ci[x][y] = rg_mt<E>::mfma(ai[x], br[y], ci[x][y]);
                }
        }
        __syncthreads();
    }
    E *Cb = C + (size_t)b * sC;
    const bool rd = beta.x != 0.0 || beta.y != 0.0;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int gr = ti0 + wr + 16 * x + rg_mt<E>::row(v, lg), gc = tj0 + wc + 16 * y + li;
                if (gr >= m || gc >= n) continue;
                const double xr = cr[x][y][v], xi = ci[x][y][v];
                double outr = alpha.x * xr - alpha.y * xi, outi = alpha.x * xi + alpha.y * xr;
                if (rd) {
                    const E o = Cb[(size_t)gr * ldc + gc];
                    outr += beta.x * o.x - beta.y * o.y;
                    outi += beta.x * o.y + beta.y * o.x;
                }
                Cb[(size_t)gr * ldc + gc] = rg_mt<E>::make(outr, outi);
            }
}

// The six block arrays of the work area for realization offset `off` (each array holds cap
// realizations of one block shape): A n1 x n1, B n1 x n2, C n2 x n1, D n2 x n2, X n1 x n2, Y n2 x n1.
template <class E>
struct Blocks {
    E *a, *b, *c, *d, *x, *y;
    size_t sa, sb, sd;  // per-realization strides: n1^2, n1 n2, n2^2
};

template <class E>
Blocks<E> blocks_of(const rg_plan *P)
{
    const size_t n1 = 128, n2 = (size_t)P->np - 128, cap = (size_t)P->batch_max;
    E *base = reinterpret_cast<E *>(P->d_schur);
    Blocks<E> w;
    w.sa = n1 * n1;
    w.sb = n1 * n2;
    w.sd = n2 * n2;
    w.a = base;
    w.b = w.a + cap * w.sa;
    w.c = w.b + cap * w.sb;
    w.d = w.c + cap * w.sb;
    w.x = w.d + cap * w.sd;
    w.y = w.x + cap * w.sb;
    const size_t off = (size_t)P->schur_off;
    w.a += off * w.sa;
    w.b += off * w.sb;
    w.c += off * w.sb;
    w.d += off * w.sd;
    w.x += off * w.sb;
    w.y += off * w.sb;
    return w;
}

// C = alpha A op(B) + beta C for complex128 128 x 128 blocks, one workgroup of 8 wavefronts per
// realization (the layout of the one-workgroup Dyson product): each operand element is read once, against
// twice by the 64 x 64 tiles of k_cgemm.  Wavefront w owns row tiles {w/2, w/2 + 4} and column tiles
// {w%2 + 2y}; k runs in chunks of 16 through a double-buffered LDS stage ([k][row] for A, [k][col] for
// B, row length 130 doubles), the next chunk loaded into registers while the current one is multiplied.
// TB: op(B) = B^T, element (k, j) at B[j ldb + k] (16 consecutive k of a column per 16 threads).
template <bool TB>
__global__ __launch_bounds__(512, 2) void k_zgemm_wg128(double2 alpha, const double2 *__restrict__ A, int lda, size_t sA,
                                                        const double2 *__restrict__ B, int ldb, size_t sB, double2 beta,
                                                        double2 *__restrict__ C, int ldc, size_t sC)
{
    constexpr int NP = 128, S = NP + 2, NK = NP / 16, LOADS = NP * 16 / 512;
    __shared__ double sa[2][2][16][S], sb[2][2][16][S];  // [buffer][re, im][k][row or col]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1, li = lane & 15, lg = lane >> 4;
    const double2 *Ab = A + (size_t)blockIdx.x * sA, *Bb = B + (size_t)blockIdx.x * sB;
    double2 ra[LOADS], rb[LOADS];
    auto load = [&](int kc) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            ra[e] = Ab[(size_t)(idx >> 4) * lda + kc * 16 + (idx & 15)];
            if constexpr (TB) rb[e] = Bb[(size_t)(idx >> 4) * ldb + kc * 16 + (idx & 15)];
            else rb[e] = Bb[(size_t)(kc * 16 + idx / NP) * ldb + idx % NP];
        }
    };
    auto store = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            sa[buf][0][idx & 15][idx >> 4] = ra[e].x;
            sa[buf][1][idx & 15][idx >> 4] = ra[e].y;
            const int kk = TB ? (idx & 15) : idx / NP, col = TB ? (idx >> 4) : idx % NP;
            sb[buf][0][kk][col] = rb[e].x;
            sb[buf][1][kk][col] = rb[e].y;
        }
    };
    typedef double d4 __attribute__((ext_vector_type(4)));
    d4 cr[2][4], ci[2][4];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            cr[x][y] = d4{0.0, 0.0, 0.0, 0.0};
            ci[x][y] = d4{0.0, 0.0, 0.0, 0.0};
        }
    load(0);
    store(0);
    __syncthreads();
    for (int kc = 0; kc < NK; kc++) {
        const int buf = kc & 1;
        if (kc + 1 < NK) load(kc + 1);
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kr = 4 * ks + lg;
            double ar[2], ai[2], nai[2], br[4], bi[4];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                ar[x] = sa[buf][0][kr][16 * (wr + 4 * x) + li];
                ai[x] = sa[buf][1][kr][16 * (wr + 4 * x) + li];
                nai[x] = -ai[x];
            }
#pragma unroll
            for (int y = 0; y < 4; y++) {
                br[y] = sb[buf][0][kr][16 * (wc + 2 * y) + li];
                bi[y] = sb[buf][1][kr][16 * (wc + 2 * y) + li];
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) {
                    cr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], cr[x][y], 0, 0, 0);
                    cr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[x], bi[y], cr[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], ci[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], ci[x][y], 0, 0, 0);
                }
        }
        if (kc + 1 < NK) store(buf ^ 1);
        __syncthreads();
    }
    double2 *Cb = C + (size_t)blockIdx.x * sC;
    const bool rd = beta.x != 0.0 || beta.y != 0.0;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                // v_mfma_f64_16x16x4_f64: register v of lane l holds D[4v + l/16][l%16]
                const size_t o = (size_t)(16 * (wr + 4 * x) + 4 * v + lg) * ldc + 16 * (wc + 2 * y) + li;
                const double xr = cr[x][y][v], xi = ci[x][y][v];
                double outr = alpha.x * xr - alpha.y * xi, outi = alpha.x * xi + alpha.y * xr;
                if (rd) {
                    const double2 c = Cb[o];
                    outr += beta.x * c.x - beta.y * c.y;
                    outi += beta.x * c.y + beta.y * c.x;
                }
                Cb[o] = cmk(outr, outi);
            }
}

// M = (diag(a) - V) - B Gss B in FP64 (as in the register-resident inversions), rounded to E and
// written as the blocks A, B, C, D; identity on the padding rows/columns >= n.
template <class E>
__global__ __launch_bounds__(256) void k_form_blocks(const E *__restrict__ gss, const rg_cqv cqv, int boundtype,
                                                     const double2 *__restrict__ adiag,
                                                     const double2 *__restrict__ bvec, int n, int np, int first,
                                                     Blocks<E> w, int cs, const int32_t *__restrict__ nxr, int ix)
{
    __shared__ double s_cq[2 * 256 + 8];
    const int b = blockIdx.x, tid = threadIdx.x;
    adiag += (size_t)b * 3 * cs;  // cs = np: constants of each realization, else shared
    bvec += (size_t)b * cs;
    if (nxr && nxr[b] == ix) adiag += cs;  // this realization's last slice: its last diagonal
    if (cqv.cq) rg_load_cq(s_cq, cqv, b, tid, 256);
    __syncthreads();
    const int n1 = 128, n2 = np - 128;
    const E *G = gss + (size_t)b * np * np;
    E *Wa = w.a + b * w.sa, *Wb = w.b + b * w.sb, *Wc = w.c + b * w.sb, *Wd = w.d + b * w.sd;
    for (int e = tid; e < np * np; e += 256) {
        const int row = e / np, col = e % np;
        double2 m;
        if (row < n && col < n) {
            m = row == col ? adiag[row] : cmk(0.0, 0.0);
            if (cqv.cq) {
                if (boundtype == RG_DIRICHLET) {
                    m.x -= s_cq[abs(row - col)] - s_cq[row + col + 2];
                } else {
                    const int dq = col - row;
                    const double2 dd = dq >= 0 ? cmk(s_cq[2 * dq], s_cq[2 * dq + 1]) : cmk(s_cq[-2 * dq], -s_cq[-2 * dq + 1]);
                    m = csub(m, dd);
                }
            }
            if (!first) {
                const E g = G[e];
                m = csub(m, cmul(cmul(bvec[row], cmk((double)g.x, (double)g.y)), bvec[col]));
            }
        } else {
            m = cmk(row == col ? 1.0 : 0.0, 0.0);
        }
        const E v = rg_mt<E>::make(m.x, m.y);
        if (row < n1) {
            if (col < n1) Wa[row * n1 + col] = v;
            else Wb[row * n2 + (col - n1)] = v;
        } else {
            if (col < n1) Wc[(row - n1) * n1 + col] = v;
            else Wd[(row - n1) * n2 + (col - n1)] = v;
        }
    }
}

// dst (m x n, ldd) = src (m x n, lds) per realization.
template <class E>
__global__ __launch_bounds__(256) void k_copy_block(const E *__restrict__ src, int lds, size_t ss, E *__restrict__ dst,
                                                    int ldd, size_t sd, int m, int n)
{
    const E *s = src + (size_t)blockIdx.y * ss;
    E *d = dst + (size_t)blockIdx.y * sd;
    for (int e = blockIdx.x * 256 + threadIdx.x; e < m * n; e += gridDim.x * 256)
        d[(size_t)(e / n) * ldd + e % n] = s[(size_t)(e / n) * lds + e % n];
}

template <class E, bool TB = false>
void cgemm(rg_plan *P, int nb, int m, int n, int k, double2 alpha, const E *A, int lda, size_t sA, const E *B, int ldb,
           size_t sB, double2 beta, E *C, int ldc, size_t sC)
{
    if constexpr (std::is_same<E, double2>::value) {
        if (m == 128 && n == 128 && k == 128) {  // complex128 blocks of order 256: one workgroup per realization
            hipLaunchKernelGGL(k_zgemm_wg128<TB>, dim3(nb), dim3(512), 0, P->stream, alpha, A, lda, sA, B, ldb, sB, beta, C,
                               ldc, sC);
            return;
        }
    }
    hipLaunchKernelGGL((k_cgemm<E, TB>), dim3((n + 63) / 64, (m + 63) / 64, nb), dim3(256), 0, P->stream, m, n, k, alpha, A,
                       lda, sA, B, ldb, sB, beta, C, ldc, sC);
}

// dst (n x m, ldd) = src (m x n, lds)^T per realization, through 32 x 32 tiles in LDS (coalesced reads and
// writes); grid (ceil(n / 32), ceil(m / 32), realizations).
template <class E>
__global__ __launch_bounds__(256) void k_copy_block_t(const E *__restrict__ src, int lds, size_t ss, E *__restrict__ dst,
                                                      int ldd, size_t sd, int m, int n)
{
    __shared__ E t[32][33];
    const E *s = src + (size_t)blockIdx.z * ss;
    E *d = dst + (size_t)blockIdx.z * sd;
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8)
        if (i0 + r < m && j0 + tx < n) t[r][tx] = s[(size_t)(i0 + r) * lds + j0 + tx];
    __syncthreads();
    for (int r = ty; r < 32; r += 8)
        if (j0 + r < n && i0 + tx < m) d[(size_t)(j0 + r) * ldd + i0 + tx] = t[tx][r];
}

// dst (m x m, ldd) = (src + src^T) / 2 per realization (src m x m, lds), the transposed tile through LDS;
// grid (ceil(m / 32), ceil(m / 32), realizations).
template <class E>
__global__ __launch_bounds__(256) void k_copy_block_sym(const E *__restrict__ src, int lds, size_t ss,
                                                        E *__restrict__ dst, int ldd, size_t sd, int m)
{
    __shared__ E t[32][33];
    const E *s = src + (size_t)blockIdx.z * ss;
    E *d = dst + (size_t)blockIdx.z * sd;
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8)  // t[r][c] = src[j0 + r][i0 + c]
        if (j0 + r < m && i0 + tx < m) t[r][tx] = s[(size_t)(j0 + r) * lds + i0 + tx];
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < m && j < m) {
            const E u = s[(size_t)i * lds + j], v = t[tx][r];  // src[i][j], src[j][i]
            d[(size_t)i * ldd + j] = rg_mt<E>::make(0.5 * ((double)u.x + (double)v.x), 0.5 * ((double)u.y + (double)v.y));
        }
    }
}

template <class E>
void copy_block(rg_plan *P, int nb, const E *src, int lds, size_t ss, E *dst, int ldd, size_t sd, int m, int n)
{
    hipLaunchKernelGGL(k_copy_block<E>, dim3((m * n + 255) / 256, nb), dim3(256), 0, P->stream, src, lds, ss, dst, ldd, sd, m,
                       n);
}

// Inverts, in place, the contiguous order-no matrices at `mats` (stride no^2) with the kernels of
// that order through a view of the plan: zero diagonal, no disorder, b = i, so that their slice
// matrix -b G b is G itself.  Pivoting stays inside the block (FP64: partial pivoting; mixed:
// verified diagonal pivots with the pivoted fallback).
template <class E>
void invert_block(const rg_plan *P, int nb, int no, E *mats, int precision)
{
    rg_plan V(*P);
    V.graphs.clear();
    for (auto &t : V.timers) { t.pending.clear(); t.pool.clear(); }
    V.profiling = false;
    V.np = no;
    V.nmu = no;
    V.cstride = 0;
    V.d_nx_r = nullptr;
    V.d_adiag = P->d_schur_zero;
    V.d_b = P->d_schur_i;
    if (precision == RG_FP64) {
        V.d_gss = reinterpret_cast<double2 *>(mats);
        if ((P->inv_fp64 == RG_INV_NOPIV || P->inv_fp64 == RG_INV_SYM) && no >= 32) {
            rg_launch_inverse_nopiv64(&V, nb, 0, 0, false, V.d_redo);
            rg_launch_inverse_gj(&V, nb, 0, 0, false, precision, V.d_redo);
        } else {
            rg_launch_inverse_f64(&V, nb, 0, 0, false);
        }
    } else {
        V.d_gss32 = reinterpret_cast<float2 *>(mats);
        if (no >= 32) {
            rg_launch_inverse_nopiv(&V, nb, 0, 0, false, V.d_redo);
            rg_launch_inverse_gj(&V, nb, 0, 0, false, precision, V.d_redo);
        } else {
            rg_launch_inverse_gj(&V, nb, 0, 0, false, precision);
        }
    }
}

// Inverts, in place, symmetric complex128 blocks of order `no` (their lower triangles) on the symmetric
// kernel through a view of the plan as invert_block does; flagged realizations go to the pivoted kernel.
void invert_block_sym(const rg_plan *P, int nb, int no, double2 *mats)
{
    rg_plan V(*P);
    V.graphs.clear();
    for (auto &t : V.timers) { t.pending.clear(); t.pool.clear(); }
    V.profiling = false;
    V.np = no;
    V.nmu = no;
    V.cstride = 0;
    V.d_nx_r = nullptr;
    V.d_adiag = P->d_schur_zero;
    V.d_b = P->d_schur_i;
    V.d_gss = mats;
    rg_launch_inverse_sym64(&V, nb, 0, 0, false, V.d_redo);
    rg_launch_inverse_gj(&V, nb, 0, 0, false, RG_FP64, V.d_redo);
}

template <class E>
void inverse_schur(rg_plan *P, int nb, int slot, int first, bool with_disorder, int precision, E *gss)
{
    const int np = P->np, n1 = 128, n2 = np - 128;
    const Blocks<E> w = blocks_of<E>(P);
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    hipLaunchKernelGGL(k_form_blocks<E>, dim3(nb), dim3(256), 0, P->stream, gss, cq, P->p.boundtype,
                       P->d_adiag + (size_t)slot * np, P->d_b, P->nmu, np, first, w, P->cstride, P->d_nx_r, P->cur_ix);
    const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
    const size_t gs = (size_t)np * np;
    if constexpr (std::is_same<E, double2>::value) {
        if (P->p.boundtype == RG_DIRICHLET && P->inv_fp64 == RG_INV_SYM && n2 >= 32) {
            invert_block_sym(P, nb, n1, w.a);                                                      // A^-1
            cgemm(P, nb, n1, n2, n1, one, w.a, n1, w.sa, w.b, n2, w.sb, zero, w.x, n2, w.sb);       // X = A^-1 B
            cgemm(P, nb, n2, n2, n1, mone, w.c, n1, w.sb, w.x, n2, w.sb, one, w.d, n2, w.sd);       // S = D - C X
            invert_block_sym(P, nb, n2, w.d);                                                      // S^-1
            cgemm(P, nb, n1, n2, n2, mone, w.x, n2, w.sb, w.d, n2, w.sd, zero, gss + n1, np, gs);   // Z = -X S^-1
            cgemm<E, true>(P, nb, n1, n1, n2, mone, gss + n1, np, gs, w.x, n2, w.sb, one, w.a, n1, w.sa);  // A^-1 - Z X^T
            hipLaunchKernelGGL(k_copy_block_sym<E>, dim3(n1 / 32, n1 / 32, nb), dim3(256), 0, P->stream, w.a, n1, w.sa,
                               gss, np, gs, n1);
            hipLaunchKernelGGL(k_copy_block_t<E>, dim3((n2 + 31) / 32, n1 / 32, nb), dim3(256), 0, P->stream, gss + n1,
                               np, gs, gss + (size_t)n1 * np, np, gs, n1, n2);                      // Z^T
            copy_block(P, nb, w.d, n2, w.sd, gss + (size_t)n1 * np + n1, np, gs, n2, n2);
            return;
        }
    }
    invert_block(P, nb, n1, w.a, precision);                                                    // A^-1
    cgemm(P, nb, n1, n2, n1, one, w.a, n1, w.sa, w.b, n2, w.sb, zero, w.x, n2, w.sb);           // X = A^-1 B
    cgemm(P, nb, n2, n2, n1, mone, w.c, n1, w.sb, w.x, n2, w.sb, one, w.d, n2, w.sd);           // S = D - C X
    cgemm(P, nb, n2, n1, n1, one, w.c, n1, w.sb, w.a, n1, w.sa, zero, w.y, n1, w.sb);           // Y = C A^-1
    invert_block(P, nb, n2, w.d, precision);                                                    // S^-1
    cgemm(P, nb, n1, n2, n2, mone, w.x, n2, w.sb, w.d, n2, w.sd, zero, gss + n1, np, gs);       // -X S^-1
    cgemm(P, nb, n2, n1, n2, mone, w.d, n2, w.sd, w.y, n1, w.sb, zero, gss + (size_t)n1 * np, np, gs);  // -S^-1 Y
    cgemm(P, nb, n1, n1, n2, mone, gss + n1, np, gs, w.y, n1, w.sb, one, w.a, n1, w.sa);        // A^-1 + X S^-1 Y
    copy_block(P, nb, w.a, n1, w.sa, gss, np, gs, n1, n1);
    copy_block(P, nb, w.d, n2, w.sd, gss + (size_t)n1 * np + n1, np, gs, n2, n2);
}

}  // namespace

size_t rg_schur_elems(int np)
{
    const size_t n1 = 128, n2 = (size_t)np - 128;
    return n1 * n1 + 4 * n1 * n2 + n2 * n2;  // A, B, C, D, X, Y (complex128 elements per realization)
}

void rg_zgemm(rg_plan *P, int nb, int m, int n, int k, double2 alpha, const double2 *A, int lda, size_t sA,
              const double2 *B, int ldb, size_t sB, double2 beta, double2 *C, int ldc, size_t sC)
{
    cgemm<double2>(P, nb, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
}

void rg_launch_inverse_schur(rg_plan *P, int nb, int slot, int first, bool with_disorder, int precision)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_INVERSE, &stop);
    if (precision == RG_FP64) inverse_schur<double2>(P, nb, slot, first, with_disorder, precision, P->d_gss);
    else inverse_schur<float2>(P, nb, slot, first, with_disorder, precision, P->d_gss32);
    rg_timer_end(P, RGK_INVERSE, stop);
}
