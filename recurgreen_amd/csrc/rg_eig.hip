// rg_eig.hip -- transmission eigenvalues: Fisher-Lee + batched Hermitian eigensolver (gfx950).
//
// src/waveguide.f90:443-449 forms t_ij = 2i sqrt(mu_i) G_ij(L;0) sqrt(mu_j) over the aperture
// ranges and returns T = sigma(t)^2 in descending order (zgesvd, src/base_utils.f90:341-380).
// Here the T are computed as the eigenvalues of the smaller Gram matrix A = t^H t (or t t^H),
// in FP64 whatever the precision of G, one workgroup per realization:
//   1. A = Y^H Y in registers (4x4 tiles per thread), Y the scaled aperture block of G, then to LDS
//      as a packed lower triangle (column-major packed, 132 KiB at n = 128); for 128 < n <= 256 the
//      packed triangle lives in a global work area instead (one per realization, through L2);
//   2. Householder reduction to a real symmetric tridiagonal (d, e) in LDS (the zhetd2 sequence:
//      v from zlarfg, p = tau A v, w = p - tau/2 (p^H v) v, A -= v w^H + w v^H);
//   3. the eigenvalues of (d, e) by Sturm-count multisection, two threads per eigenvalue, each
//      one found by its index, so the output is sorted without a sort.
// Eigenvalues of A have absolute accuracy ~ n eps ||A||; values that come out below zero (only
// round-off of an exactly or nearly singular t) are returned as 0.
#include "rg_internal.h"

namespace {

constexpr int EG_NT = 512;  // threads per workgroup
constexpr int EG_MAXN = 128;   // largest order with the packed triangle in LDS
constexpr int EG_MAXN2 = 256;  // largest order (packed triangle in global memory)

__device__ inline double2 ldg2(const double2 *p) { return *p; }
__device__ inline double2 ldg2(const float2 *p)
{
    const float2 v = *p;
    return cmk((double)v.x, (double)v.y);
}

__device__ inline int pk_off(int j, int n) { return j * n - (j * (j - 1)) / 2; }   // start of packed column j
__device__ inline int pk(int i, int j, int n) { return pk_off(j, n) + (i - j); }  // i >= j

// 1/q to full FP64 accuracy: hardware reciprocal and two Newton steps.
__device__ inline double rcp_f64(double q)
{
    double r = __builtin_amdgcn_rcp(q);
    r = fma(fma(-q, r, 1.0), r, r);
    r = fma(fma(-q, r, 1.0), r, r);
    return r;
}

// Number of eigenvalues of the tridiagonal (d, e) below x (Sturm sequence of the LDL^T pivots).
__device__ inline int sturm_count(const double *d, const double *e2, int n, double x, double pivmin)
{
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < n; i++) {
        q = (d[i] - x) - e2[i - 1] * rcp_f64(q);
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// VEC = false: the neigv largest eigenvalues to teig.  VEC = true: only the top eigenpair; its
// eigenvector (unit norm, n complex) to xvec[b * EG_MAXN ..]: inverse iteration on the tridiagonal
// (partial-pivoting LU, as LAPACK's dlagtf/dlagts) and back-transformation by the Householder
// reflectors, which are kept below the subdiagonal of the packed triangle.
// MAXN = EG_MAXN keeps the packed triangle in LDS (132 KiB); MAXN = EG_MAXN2 keeps it in the global
// work area `work` (one packed triangle of MAXN per realization, streamed through L2).
template <class E, bool VEC, int MAXN>
__global__ __launch_bounds__(EG_NT, 1) void k_tdt_eig(const E *__restrict__ gs1, int NP, int imin, int imax, int jmin,
                                                      int jmax, const double2 *__restrict__ sqrtmu,
                                                      double *__restrict__ teig, int neigv,
                                                      int32_t *__restrict__ status, double2 *__restrict__ xvec, int cs,
                                                      double2 *__restrict__ work)
{
    sqrtmu += (size_t)blockIdx.x * cs;  // cs = np: constants of each realization, else shared
    __shared__ double2 sA_lds[MAXN == EG_MAXN ? EG_MAXN * (EG_MAXN + 1) / 2 : 1];
    double2 *sA = MAXN == EG_MAXN ? sA_lds : work + (size_t)blockIdx.x * NP * NP;  // n (n + 1) / 2 <= NP^2
    __shared__ double2 s_taus[VEC ? MAXN : 1];
    __shared__ double2 s_x[VEC ? MAXN : 1];
    __shared__ double s_lam;
    __shared__ double2 s_v[MAXN], s_p[MAXN];
    __shared__ double s_d[MAXN], s_e[MAXN], s_e2[MAXN];
    __shared__ double s_w[MAXN];     // |sqrt(mu_l)|^2 weights
    __shared__ double2 s_sa[MAXN];  // column scales of A
    __shared__ double2 s_tau;
    __shared__ int s_bad;

    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const bool gram_cols = nn <= m;  // A = t^H t (n = nn), else A = t t^H (n = m)
    const int n = gram_cols ? nn : m, L = gram_cols ? m : nn;
    const E *G = gs1 + (size_t)b * NP * NP;

    // A_ab = 4 conj(s_a) s_b sum_l w_l conj(G_la) G_lb with, for t^H t: G_la = G[il][ja], w_l = |mu_il|,
    // s_a = sqrt(mu_ja); for t t^H: G_la = conj(G[ia][jl]), w_l = |mu_jl|, s_a = conj(sqrt(mu_ia)).
    if (tid == 0) s_bad = 0;
    for (int l = tid; l < L; l += EG_NT) {
        const double2 s = sqrtmu[(gram_cols ? imin : jmin) - 1 + l];
        s_w[l] = s.x * s.x + s.y * s.y;
    }
    for (int a = tid; a < n; a += EG_NT) {
        const double2 s = sqrtmu[(gram_cols ? jmin : imin) - 1 + a];
        s_sa[a] = gram_cols ? s : cconj(s);
    }
    __syncthreads();

    // 1. Gram matrix, 4x4 tiles of the lower triangle.
    const int nt4 = (n + 3) / 4, ntiles = nt4 * (nt4 + 1) / 2;
    for (int tile = tid; tile < ntiles; tile += EG_NT) {
        // tile -> (ta >= tb): ta = largest with ta (ta+1) / 2 <= tile
        int ta = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
        while ((ta + 1) * (ta + 2) / 2 <= tile) ta++;
        while (ta * (ta + 1) / 2 > tile) ta--;
        const int tb = tile - ta * (ta + 1) / 2;
        double2 acc[4][4];
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) acc[x][y] = cmk(0.0, 0.0);
        for (int l = 0; l < L; l++) {
            double2 ga[4], gb[4];
            const double wl = s_w[l];
#pragma unroll
            for (int x = 0; x < 4; x++) {
                const int ia = 4 * ta + x, ib = 4 * tb + x;
                double2 va = cmk(0.0, 0.0), vb = cmk(0.0, 0.0);
                if (gram_cols) {
                    if (ia < n) va = ldg2(G + (size_t)(imin - 1 + l) * NP + (jmin - 1 + ia));
                    if (ib < n) vb = ldg2(G + (size_t)(imin - 1 + l) * NP + (jmin - 1 + ib));
                } else {
                    if (ia < n) va = cconj(ldg2(G + (size_t)(imin - 1 + ia) * NP + (jmin - 1 + l)));
                    if (ib < n) vb = cconj(ldg2(G + (size_t)(imin - 1 + ib) * NP + (jmin - 1 + l)));
                }
                ga[x] = cmk(wl * va.x, -wl * va.y);  // w_l conj(G_la)
                gb[x] = vb;
            }
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) acc[x][y] = cfma(acc[x][y], ga[x], gb[y]);
        }
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) {
                const int ia = 4 * ta + x, ib = 4 * tb + y;
                if (ia < n && ib < n && ia >= ib) {
                    const double2 sc = cmul(cconj(s_sa[ia]), s_sa[ib]);
                    const double2 v = cmul(sc, acc[x][y]);
                    sA[pk(ia, ib, n)] = cmk(4.0 * v.x, 4.0 * v.y);
                }
            }
    }
    __syncthreads();

    // 2. Householder tridiagonalisation of the packed lower triangle.
    for (int k = 0; k + 1 < n; k++) {
        const int Lk = n - k - 1;  // order of the trailing block
        double2 *xk = sA + pk(k + 1, k, n);
        if (tid < 64) {
            double ss = 0.0;
            for (int t = 1 + lane; t < Lk; t += 64) ss += xk[t].x * xk[t].x + xk[t].y * xk[t].y;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) ss += __shfl_xor(ss, off, 64);
            const double2 alpha = xk[0];
            double2 tau = cmk(0.0, 0.0), scale = cmk(0.0, 0.0);
            double beta = alpha.x;
            if (ss > 0.0 || alpha.y != 0.0) {
                const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + ss);
                beta = alpha.x >= 0.0 ? -nrm : nrm;
                tau = cmk((beta - alpha.x) / beta, -alpha.y / beta);
                scale = crecip(cmk(alpha.x - beta, alpha.y));
            }
            for (int t = lane; t < Lk; t += 64) {
                const double2 vt = t == 0 ? cmk(1.0, 0.0) : cmul(xk[t], scale);
                s_v[t] = vt;
                if (VEC && t > 0) xk[t] = vt;  // column k below the subdiagonal is not read again
            }
            if (lane == 0) {
                if (VEC) s_taus[k] = tau;
                s_tau = tau;
                s_d[k] = sA[pk(k, k, n)].x;
                s_e[k] = beta;
            }
        }
        __syncthreads();
        const double2 tau = s_tau;
        if (tau.x == 0.0 && tau.y == 0.0) {  // H = I; uniform, and s_tau is not rewritten before this barrier
            __syncthreads();
            continue;
        }
        // p = tau A22 v; four threads per row, each a quarter of the columns (rows in passes of
        // EG_NT / 4, the same count for every lane)
        for (int r0 = 0; r0 < Lk; r0 += EG_NT / 4) {
            const int r = r0 + (tid >> 2), h = tid & 3;
            double2 acc = cmk(0.0, 0.0);
            if (r < Lk) {
                const int c0 = (h * Lk) >> 2, c1 = ((h + 1) * Lk) >> 2;
                const int gr = k + 1 + r;
                for (int c = c0; c < c1; c++) {
                    const int gc = k + 1 + c;
                    const double2 a = c <= r ? sA[pk(gr, gc, n)] : cconj(sA[pk(gc, gr, n)]);
                    acc = cfma(acc, a, s_v[c]);
                }
            }
            acc.x += __shfl_xor(acc.x, 1, 64);
            acc.y += __shfl_xor(acc.y, 1, 64);
            acc.x += __shfl_xor(acc.x, 2, 64);
            acc.y += __shfl_xor(acc.y, 2, 64);
            if (r < Lk && h == 0) s_p[r] = cmul(tau, acc);
        }
        __syncthreads();
        // alpha2 = -tau/2 (p^H v), by every wavefront; w = p + alpha2 v formed on the fly
        double2 dot = cmk(0.0, 0.0);
        for (int t = lane; t < Lk; t += 64) dot = cfma(dot, cconj(s_p[t]), s_v[t]);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            dot.x += __shfl_xor(dot.x, off, 64);
            dot.y += __shfl_xor(dot.y, off, 64);
        }
        const double2 alpha2 = cmul(cmk(-0.5 * tau.x, -0.5 * tau.y), dot);
        // A22 -= v w^H + w v^H on the lower triangle: column pairs (c, Lk-1-c), eight threads a pair
        for (int pr0 = 0, npair = (Lk + 1) >> 1; pr0 < npair; pr0 += EG_NT / 8) {
            const int pr = pr0 + (tid >> 3), sub = tid & 7;
            if (pr < npair) {
                const int c1 = pr, c2 = Lk - 1 - pr;
                const int len1 = Lk - c1, len = c1 == c2 ? len1 : Lk + 1;
                const double2 v1 = s_v[c1], w1 = cfma(s_p[c1], alpha2, v1);
                const double2 v2 = s_v[c2], w2 = cfma(s_p[c2], alpha2, v2);
                for (int s = sub; s < len; s += 8) {
                    const bool first = s < len1;
                    const int c = first ? c1 : c2;
                    const int r = first ? c1 + s : c2 + (s - len1);
                    const double2 vc = first ? v1 : v2, wc = first ? w1 : w2;
                    const double2 vr = s_v[r], wr = cfma(s_p[r], alpha2, vr);
                    double2 *ap = sA + pk(k + 1 + r, k + 1 + c, n);
                    double2 a = *ap;
                    a = csub(a, cmul(vr, cconj(wc)));
                    a = csub(a, cmul(wr, cconj(vc)));
                    *ap = a;
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        s_d[n - 1] = sA[pk(n - 1, n - 1, n)].x;
        s_e[n - 1] = 0.0;
    }
    __syncthreads();
    for (int i = tid; i < n; i += EG_NT) {
        s_e2[i] = s_e[i] * s_e[i];
        if (!isfinite(s_d[i]) || !isfinite(s_e[i])) s_bad = 1;
    }
    __syncthreads();
    if (s_bad) {
        if (tid == 0) status[b] = RG_ENONFINITE;
        return;
    }

    // 3. Eigenvalues by Sturm-count trisection: thread pair (j, h) finds ascending eigenvalue j.
    double lo = 1e300, hi = -1e300, e2max = 0.0;
    for (int i = 0; i < n; i++) {
        const double ri = (i > 0 ? fabs(s_e[i - 1]) : 0.0) + (i + 1 < n ? fabs(s_e[i]) : 0.0);
        lo = fmin(lo, s_d[i] - ri);
        hi = fmax(hi, s_d[i] + ri);
        e2max = fmax(e2max, s_e2[i]);
    }
    const double span = fmax(hi - lo, 1e-300);
    lo -= 2.220446049250313e-16 * span * n;
    hi += 2.220446049250313e-16 * span * n;
    const double pivmin = 1e-290 * fmax(1.0, e2max);
    const int j = tid >> 1, h = tid & 1;
    if (j < n && (VEC ? j == n - 1 : n - 1 - j < neigv)) {
        double a = lo, z = hi;
        for (int it = 0; it < 40; it++) {
            const double step = (z - a) * (1.0 / 3.0);
            const double x = h == 0 ? a + step : z - step;
            const int cnt = sturm_count(s_d, s_e2, n, x, pivmin);
            const int other = __shfl_xor(cnt, 1, 64);
            const int c1 = h == 0 ? cnt : other, c2 = h == 0 ? other : cnt;
            const double x1 = a + step, x2 = z - step;
            if (c1 > j) z = x1;
            else if (c2 > j) { a = x1; z = x2; }
            else a = x2;
            if (z - a <= 2.220446049250313e-16 * fmax(fabs(a), fabs(z)) + pivmin) break;
        }
        if (h == 0) {
            if (VEC) s_lam = 0.5 * (a + z);
            else {
                // complex64 state: T projected onto [0, 1], where every exact eigenvalue lies
                // (unitarity, no gain), so FP32 round-off does not push open channels out of range
                const double tv = fmax(0.5 * (a + z), 0.0);
                teig[(size_t)b * neigv + (n - 1 - j)] = std::is_same<E, float2>::value ? fmin(tv, 1.0) : tv;
            }
        }
    }
    if constexpr (VEC) {
        __syncthreads();
        if (tid == 0) {
            // (T - sigma I) z = r by Gaussian elimination with row interchanges; rows kept as
            // (u0, u1, u2) = entries at columns i, i+1, i+2.  Three solves from r = 1.
            double *u0 = reinterpret_cast<double *>(s_p), *u1 = u0 + MAXN, *u2 = s_e2;
            double *zz = reinterpret_cast<double *>(s_v), *rh = zz + MAXN;
            const double sigma = s_lam + 1e-14 * fmax(1.0, fabs(s_lam));
            for (int i = 0; i < n; i++) zz[i] = 1.0;
            for (int it = 0; it < 3; it++) {
                for (int i = 0; i < n; i++) {
                    u0[i] = s_d[i] - sigma;
                    u1[i] = i + 1 < n ? s_e[i] : 0.0;
                    u2[i] = 0.0;
                    rh[i] = zz[i];
                }
                for (int i = 0; i + 1 < n; i++) {
                    const double sub = s_e[i];  // row i+1, column i
                    if (fabs(sub) > fabs(u0[i])) {
                        const double r0 = u0[i], r1 = u1[i], r2 = u2[i];
                        u0[i] = sub;
                        u1[i] = u0[i + 1];
                        u2[i] = u1[i + 1];
                        const double mlt = r0 / sub;
                        u0[i + 1] = r1 - mlt * u1[i];
                        u1[i + 1] = r2 - mlt * u2[i];
                        const double t = rh[i];
                        rh[i] = rh[i + 1];
                        rh[i + 1] = t - mlt * rh[i];
                    } else {
                        if (u0[i] == 0.0) u0[i] = pivmin;
                        const double mlt = sub / u0[i];
                        u0[i + 1] -= mlt * u1[i];
                        u1[i + 1] -= mlt * u2[i];
                        rh[i + 1] -= mlt * rh[i];
                    }
                }
                if (u0[n - 1] == 0.0) u0[n - 1] = pivmin;
                double nrm = 0.0;
                for (int i = n - 1; i >= 0; i--) {
                    double v = rh[i];
                    if (i + 1 < n) v -= u1[i] * zz[i + 1];
                    if (i + 2 < n) v -= u2[i] * zz[i + 2];
                    zz[i] = v / u0[i];
                    nrm += zz[i] * zz[i];
                }
                const double inv = 1.0 / sqrt(nrm);
                for (int i = 0; i < n; i++) zz[i] *= inv;
            }
            for (int i = 0; i < n; i++) s_x[i] = cmk(zz[i], 0.0);
        }
        __syncthreads();
        // x <- H_0 H_1 ... H_{n-2} z, H_k = I - tau_k v_k v_k^H on rows k+1.. (v_k[0] = 1)
        if (tid < 64) {
            for (int k = n - 2; k >= 0; k--) {
                const double2 tau = s_taus[k];
                if (tau.x == 0.0 && tau.y == 0.0) continue;
                const int Lk = n - k - 1;
                const double2 *vk = sA + pk(k + 1, k, n);
                double2 dot = cmk(0.0, 0.0);
                for (int t = lane; t < Lk; t += 64) {
                    const double2 vt = t == 0 ? cmk(1.0, 0.0) : vk[t];
                    dot = cfma(dot, cconj(vt), s_x[k + 1 + t]);
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    dot.x += __shfl_xor(dot.x, off, 64);
                    dot.y += __shfl_xor(dot.y, off, 64);
                }
                const double2 f = cmul(tau, dot);
                for (int t = lane; t < Lk; t += 64) {
                    const double2 vt = t == 0 ? cmk(1.0, 0.0) : vk[t];
                    s_x[k + 1 + t] = csub(s_x[k + 1 + t], cmul(vt, f));
                }
            }
            for (int i = lane; i < n; i += 64) xvec[(size_t)b * EG_MAXN + i] = s_x[i];
        }
    }
}

}  // namespace

bool rg_topvec_supported(const rg_plan *P)
{
    const int m = P->imax - P->imin + 1, nn = P->jmax - P->jmin + 1;
    return (m < nn ? m : nn) <= EG_MAXN;
}

bool rg_eig_supported(const rg_plan *P)
{
    const int m = P->imax - P->imin + 1, nn = P->jmax - P->jmin + 1;
    return (m < nn ? m : nn) <= EG_MAXN2;
}

// Orders up to EG_MAXN in LDS; up to EG_MAXN2 with the packed triangle in the t buffer (free while
// eigenvalues are computed: the sweep's state is in Gss/Gs1).
template <class E>
static void launch_eig(rg_plan *P, int nb, const E *gs1, double *teig, int32_t *status)
{
    const int m = P->imax - P->imin + 1, nn = P->jmax - P->jmin + 1;
    if ((m < nn ? m : nn) <= EG_MAXN)
        hipLaunchKernelGGL((k_tdt_eig<E, false, EG_MAXN>), dim3(nb), dim3(EG_NT), 0, P->stream, gs1, P->np, P->imin,
                           P->imax, P->jmin, P->jmax, P->d_sqrtmu, teig, P->neigv, status, nullptr, P->cstride, nullptr);
    else
        hipLaunchKernelGGL((k_tdt_eig<E, false, EG_MAXN2>), dim3(nb), dim3(EG_NT), 0, P->stream, gs1, P->np, P->imin,
                           P->imax, P->jmin, P->jmax, P->d_sqrtmu, teig, P->neigv, status, nullptr, P->cstride,
                           P->d_tmat);
}

// Order <= EG_MAXN only (the adjoint's entry points refuse np > 128).
void rg_launch_tdt_topvec_f64(rg_plan *P, int nb, const double2 *gs1, int32_t *status, double2 *xvec)
{
    hipLaunchKernelGGL((k_tdt_eig<double2, true, EG_MAXN>), dim3(nb), dim3(EG_NT), 0, P->stream, gs1, P->np, P->imin,
                       P->imax, P->jmin, P->jmax, P->d_sqrtmu, nullptr, P->neigv, status, xvec, P->cstride, nullptr);
}

void rg_launch_tdt_topvec_f32(rg_plan *P, int nb, const float2 *gs1, int32_t *status, double2 *xvec)
{
    hipLaunchKernelGGL((k_tdt_eig<float2, true, EG_MAXN>), dim3(nb), dim3(EG_NT), 0, P->stream, gs1, P->np, P->imin,
                       P->imax, P->jmin, P->jmax, P->d_sqrtmu, nullptr, P->neigv, status, xvec, P->cstride, nullptr);
}

// Orders beyond EG_MAXN2 take the one-sided Jacobi SVD, so no launch indexes past the arrays.
void rg_launch_tdt_eig_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status)
{
    if (!rg_eig_supported(P)) return rg_launch_fisher_lee_svd_f64(P, nb, gs1, teig, status, nullptr);
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    launch_eig<double2>(P, nb, gs1, teig, status);
    rg_timer_end(P, RGK_SVD, stop);
}

void rg_launch_tdt_eig_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status)
{
    if (!rg_eig_supported(P)) return rg_launch_fisher_lee_svd_f32(P, nb, gs1, teig, status, nullptr);
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    launch_eig<float2>(P, nb, gs1, teig, status);
    rg_timer_end(P, RGK_SVD, stop);
}
