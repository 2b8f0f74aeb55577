// rg_tprofile.hip -- transmission eigenvectors and eigenstate intensities for the tprofile tasks (gfx950).
//
// compute_tprofile / compute_tprofile_group (src/waveguide.f90:843-997) need, per realization, the
// eigen-decomposition t^H t = U diag(T) U^H of the transmission matrix t = 2i s G(L;0) s,
// s = sqrt(Re mu), over the rows imin.. and columns jmin.. of a range (the full nmu x nmu matrix for
// compute_tprofile, the apertures for the group variant; diagonalize_ahca, src/base_utils.f90:394-443:
// T = sigma^2, U the right singular vectors), and the intensities |G(x;1) s U(:,v)|^2 summed over
// the modes at every slice x.  k_tprofile_svd is a one-sided (Hestenes) Jacobi SVD of t whose
// rotations are also applied to V = I, so that t V has orthogonal columns: V = U and
// T_v = |(t V)(:,v)|^2.  It writes T in descending order and W = s U with its columns in that order
// and its rows at the column range of t (row-major np x np, the layout of the batched products, zero
// elsewhere).  k_colnorms writes the per-column norms |Y(:,v)|^2 of Y = G(x;1) W.
#include "rg_internal.h"

namespace {

constexpr int TP_MAXN = 256;  // order supported by the rank table in LDS

// Round-robin pairing of cols_e (even) columns: round r, pair pr -> (p, q), every pair once per
// sweep of cols_e - 1 rounds.
__device__ inline void tp_pair(int r, int pr, int last, int &p, int &q)
{
    if (pr == 0) {
        p = r;
        q = last;
    } else {
        p = (r + pr) % last;
        q = (r - pr + last) % last;
    }
}

__global__ __launch_bounds__(256) void k_tprofile_svd(const double2 *__restrict__ gs1, int NP, int n, int i0,
                                                      int m, int j0, int nn, const double *__restrict__ sre,
                                                      double2 *__restrict__ xwork,
                                                      double2 *__restrict__ vwork, double2 *__restrict__ w,
                                                      double *__restrict__ tvals, int32_t *__restrict__ status,
                                                      int max_sweeps)
{
    __shared__ int s_rot;
    __shared__ double s_t[TP_MAXN];
    __shared__ int s_rank[TP_MAXN];
    // t: m rows from i0, nn columns from j0 (0-based); column j of t at X + j NP, of V at V + j NP
    const int b = blockIdx.x, tid = threadIdx.x;
    const int cols_e = nn + (nn & 1);
    const double2 *G = gs1 + (size_t)b * NP * NP;
    double2 *X = xwork + (size_t)b * NP * NP;
    double2 *V = vwork + (size_t)b * NP * NP;
    double2 *Wb = w + (size_t)b * NP * NP;
    for (int e = tid; e < cols_e * NP; e += blockDim.x) {
        const int j = e / NP, i = e % NP;
        double2 t = cmk(0.0, 0.0);
        if (i < m && j < nn) {
            const double s = sre[i0 + i] * sre[j0 + j];
            const double2 g = G[(size_t)(i0 + i) * NP + j0 + j];
            t = cmk(-2.0 * s * g.y, 2.0 * s * g.x);  // 2i s g s
        }
        X[(size_t)j * NP + i] = t;
        if (i < cols_e) V[(size_t)j * NP + i] = cmk(i == j ? 1.0 : 0.0, 0.0);
    }
    for (int e = tid; e < n * n; e += blockDim.x) Wb[(size_t)(e / n) * NP + e % n] = cmk(0.0, 0.0);
    __syncthreads();

    const double tol = 2.220446049250313e-16 * sqrt((double)m);
    const int grp = tid >> 4, lane = tid & 15, ngrp = blockDim.x >> 4;
    const int last = cols_e - 1;
    int sweep = 0;
    for (; sweep < max_sweeps; sweep++) {
        if (tid == 0) s_rot = 0;
        __syncthreads();
        for (int round = 0; round < last; round++) {
            for (int pr = grp; pr < cols_e / 2; pr += ngrp) {
                int p, q;
                tp_pair(round, pr, last, p, q);
                double2 *ap = X + (size_t)p * NP, *aq = X + (size_t)q * NP;
                double al = 0.0, be = 0.0, gx = 0.0, gy = 0.0;
                for (int i = lane; i < m; i += 16) {
                    const double2 x = ap[i], y = aq[i];
                    al = fma(x.x, x.x, fma(x.y, x.y, al));
                    be = fma(y.x, y.x, fma(y.y, y.y, be));
                    gx = fma(x.x, y.x, fma(x.y, y.y, gx));  // conj(x) y
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
                const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                const double2 e = cmk(gx / g, -gy / g);  // exp(-i arg gamma)
                // [x_p, x_q] <- [x_p, x_q] [[c, sn], [-sn e, c e]], the same on the columns of V
                for (int i = lane; i < m; i += 16) {
                    const double2 x = ap[i], y = cmul(e, aq[i]);
                    ap[i] = cmk(c * x.x - sn * y.x, c * x.y - sn * y.y);
                    aq[i] = cmk(sn * x.x + c * y.x, sn * x.y + c * y.y);
                }
                double2 *vp = V + (size_t)p * NP, *vq = V + (size_t)q * NP;
                for (int i = lane; i < cols_e; i += 16) {
                    const double2 x = vp[i], y = cmul(e, vq[i]);
                    vp[i] = cmk(c * x.x - sn * y.x, c * x.y - sn * y.y);
                    vq[i] = cmk(sn * x.x + c * y.x, sn * x.y + c * y.y);
                }
            }
            __syncthreads();
        }
        if (!s_rot) break;
        __syncthreads();
    }
    for (int j = grp; j < nn; j += ngrp) {
        double s2 = 0.0;
        const double2 *aj = X + (size_t)j * NP;
        for (int i = lane; i < m; i += 16) s2 = fma(aj[i].x, aj[i].x, fma(aj[i].y, aj[i].y, s2));
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) s2 += __shfl_xor(s2, off, 16);
        if (lane == 0) s_t[j] = s2;
    }
    __syncthreads();
    bool bad = sweep == max_sweeps;
    for (int j = tid; j < nn; j += blockDim.x) {
        const double v = s_t[j];
        int rank = 0;
        for (int k2 = 0; k2 < nn; k2++) {
            const double u = s_t[k2];
            rank += (u > v) || (u == v && k2 < j);
        }
        s_rank[j] = rank;
        tvals[(size_t)b * nn + rank] = v;
        if (!(v == v) || v > 1e300) bad = true;
    }
    if (bad) status[b] = RG_ENOCONV;
    __syncthreads();
    // W = s U, column rank(j) <- s * V(:, j) on rows j0 .. j0+nn-1; rows and columns >= n are not
    // read by the products
    for (int e = tid; e < nn * nn; e += blockDim.x) {
        const int i = e / nn, j = e % nn;
        const double2 v = V[(size_t)j * NP + i];
        Wb[(size_t)(j0 + i) * NP + s_rank[j]] = cmk(sre[j0 + i] * v.x, sre[j0 + i] * v.y);
    }
}

// out[b * ostride + v] = sum_{i < n} |Y_b[i][v]|^2 for v < nv.
__global__ __launch_bounds__(256) void k_colnorms(const double2 *__restrict__ y, int NP, int n, int nv,
                                                  double *__restrict__ out, size_t ostride)
{
    const int b = blockIdx.x;
    const double2 *Y = y + (size_t)b * NP * NP;
    for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < n; i++) {
            const double2 z = Y[(size_t)i * NP + v];
            s = fma(z.x, z.x, fma(z.y, z.y, s));
        }
        out[(size_t)b * ostride + v] = s;
    }
}

// W = e_row in column 0 (the active n x n block, zero elsewhere): the right-hand side of one column of
// the Green function.
__global__ __launch_bounds__(256) void k_unit_column(double2 *__restrict__ w, int NP, int n, int row)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
        const int i = e / n, j = e % n;
        w[(size_t)i * NP + j] = cmk(i == row && j == 0 ? 1.0 : 0.0, 0.0);
    }
}

// psi[iy] = sum_m chi_m(y_iy) Y[m][0]: the transverse field of one slice from its modal column, with the
// modes of src/waveguide.f90:151-229 on y_iy = (iy + 1/2)/ny (0-based iy): sqrt(2/ny) sin(pi m y)
// (Dirichlet, m = 1..nmu) or exp(2 pi i (m - nmu/2 - 1) y)/sqrt(ny) (periodic).
__global__ __launch_bounds__(256) void k_modal_column(const double2 *__restrict__ y, int NP, int n, int ny,
                                                      int boundtype, double2 *__restrict__ psi)
{
    const int iy = blockIdx.x * blockDim.x + threadIdx.x;
    if (iy >= ny) return;
    const double yv = (iy + 0.5) / ny;
    double2 acc = cmk(0.0, 0.0);
    if (boundtype == RG_DIRICHLET) {
        const double nrm = sqrt(2.0 / ny);
        for (int m = 0; m < n; m++) {
            const double c = nrm * sinpi((m + 1) * yv);
            const double2 v = y[(size_t)m * NP];
            acc = cmk(fma(c, v.x, acc.x), fma(c, v.y, acc.y));
        }
    } else {
        const double nrm = 1.0 / sqrt((double)ny);
        for (int m = 0; m < n; m++) {
            double sn, cs;
            sincospi(2.0 * (m - n / 2) * yv, &sn, &cs);
            acc = cadd(acc, cmul(cmk(nrm * cs, nrm * sn), y[(size_t)m * NP]));
        }
    }
    psi[iy] = acc;
}

// Top eigenvector of the Gram matrix of t = 2i s G s over the filtered ranges with the complex
// sqrt(mu) of src/waveguide.f90:443 (the adjoint of T_max): t^H t when t has no more columns than rows,
// else t t^H, as the eigenvalue kernel orients it.  One-sided Jacobi on the columns of t (or of t^H)
// with the rotations accumulated into V; the column of largest norm gives the eigenvector, written
// (unit norm, n complex) to xvec[b * xstride ..].  For orders the LDS-resident eigensolver does not hold.
template <class E>
__global__ __launch_bounds__(256) void k_topvec_jacobi(const E *__restrict__ gs1, int NP, int imin, int imax, int jmin,
                                                       int jmax, const double2 *__restrict__ sqrtmu,
                                                       double2 *__restrict__ xwork, double2 *__restrict__ vwork,
                                                       double2 *__restrict__ xvec, int xstride,
                                                       int32_t *__restrict__ status, int max_sweeps)
{
    __shared__ int s_rot;
    __shared__ double s_t[TP_MAXN];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int mr = imax - imin + 1, mc = jmax - jmin + 1;
    const bool gram_cols = mc <= mr;
    const int n = gram_cols ? mc : mr, rows = gram_cols ? mr : mc;  // columns of the decomposed matrix
    const int cols_e = n + (n & 1);
    const E *G = gs1 + (size_t)b * NP * NP;
    double2 *X = xwork + (size_t)b * NP * NP;
    double2 *V = vwork + (size_t)b * NP * NP;
    for (int e = tid; e < cols_e * NP; e += blockDim.x) {
        const int j = e / NP, i = e % NP;
        double2 t = cmk(0.0, 0.0);
        if (i < rows && j < n) {
            const int a = gram_cols ? i : j, c = gram_cols ? j : i;  // element (a, c) of t
            const int ii = imin - 1 + a, jj = jmin - 1 + c;
            const double2 v = cmul(cmul(sqrtmu[ii], rg_mt<E>::ld(G[(size_t)ii * NP + jj])), sqrtmu[jj]);
            t = cmk(-2.0 * v.y, 2.0 * v.x);
            if (!gram_cols) t = cconj(t);  // columns of t^H
        }
        X[(size_t)j * NP + i] = t;
        if (i < cols_e) V[(size_t)j * NP + i] = cmk(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    const double tol = 2.220446049250313e-16 * sqrt((double)rows);
    const int grp = tid >> 4, lane = tid & 15, ngrp = blockDim.x >> 4;
    const int last = cols_e - 1;
    int sweep = 0;
    for (; sweep < max_sweeps; sweep++) {
        if (tid == 0) s_rot = 0;
        __syncthreads();
        for (int round = 0; round < last; round++) {
            for (int pr = grp; pr < cols_e / 2; pr += ngrp) {
                int p, q;
                tp_pair(round, pr, last, p, q);
                double2 *ap = X + (size_t)p * NP, *aq = X + (size_t)q * NP;
                double al = 0.0, be = 0.0, gx = 0.0, gy = 0.0;
                for (int i = lane; i < rows; i += 16) {
                    const double2 x = ap[i], y = aq[i];
                    al = fma(x.x, x.x, fma(x.y, x.y, al));
                    be = fma(y.x, y.x, fma(y.y, y.y, be));
                    gx = fma(x.x, y.x, fma(x.y, y.y, gx));
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
                const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                const double2 e = cmk(gx / g, -gy / g);
                for (int i = lane; i < rows; i += 16) {
                    const double2 x = ap[i], y = cmul(e, aq[i]);
                    ap[i] = cmk(c * x.x - sn * y.x, c * x.y - sn * y.y);
                    aq[i] = cmk(sn * x.x + c * y.x, sn * x.y + c * y.y);
                }
                double2 *vp = V + (size_t)p * NP, *vq = V + (size_t)q * NP;
                for (int i = lane; i < cols_e; i += 16) {
                    const double2 x = vp[i], y = cmul(e, vq[i]);
                    vp[i] = cmk(c * x.x - sn * y.x, c * x.y - sn * y.y);
                    vq[i] = cmk(sn * x.x + c * y.x, sn * x.y + c * y.y);
                }
            }
            __syncthreads();
        }
        if (!s_rot) break;
        __syncthreads();
    }
    for (int j = grp; j < n; j += ngrp) {
        double s2 = 0.0;
        const double2 *aj = X + (size_t)j * NP;
        for (int i = lane; i < rows; i += 16) s2 = fma(aj[i].x, aj[i].x, fma(aj[i].y, aj[i].y, s2));
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) s2 += __shfl_xor(s2, off, 16);
        if (lane == 0) s_t[j] = s2;
    }
    __syncthreads();
    int top = 0;  // every thread finds the same largest column (lowest index on ties)
    for (int j = 1; j < n; j++)
        if (s_t[j] > s_t[top]) top = j;
    for (int i = tid; i < n; i += blockDim.x) xvec[(size_t)b * xstride + i] = V[(size_t)top * NP + i];
    if (tid == 0 && (sweep == max_sweeps || !(s_t[top] == s_t[top]))) status[b] = RG_ENOCONV;
}

}  // namespace

bool rg_tprofile_supported(int n) { return n <= TP_MAXN; }

template <class E>
static void launch_topvec_jacobi(rg_plan *P, int nb, const E *gs1, double2 *xwork, double2 *vwork, double2 *xvec,
                                 int xstride)
{
    hipLaunchKernelGGL(k_topvec_jacobi<E>, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->imin, P->imax, P->jmin,
                       P->jmax, P->d_sqrtmu, xwork, vwork, xvec, xstride, P->d_status, 60);
}

void rg_launch_topvec_jacobi(rg_plan *P, int nb, const double2 *gs1, double2 *xwork, double2 *vwork, double2 *xvec,
                             int xstride)
{
    launch_topvec_jacobi(P, nb, gs1, xwork, vwork, xvec, xstride);
}

void rg_launch_topvec_jacobi(rg_plan *P, int nb, const float2 *gs1, double2 *xwork, double2 *vwork, double2 *xvec,
                             int xstride)
{
    launch_topvec_jacobi(P, nb, gs1, xwork, vwork, xvec, xstride);
}

void rg_launch_unit_column(rg_plan *P, double2 *w, int row)
{
    hipLaunchKernelGGL(k_unit_column, dim3((P->nmu * P->nmu + 255) / 256), dim3(256), 0, P->stream, w, P->np, P->nmu,
                       row);
}

void rg_launch_modal_column(rg_plan *P, const double2 *y, double2 *psi)
{
    hipLaunchKernelGGL(k_modal_column, dim3((P->ny + 255) / 256), dim3(256), 0, P->stream, y, P->np, P->nmu, P->ny,
                       P->p.boundtype, psi);
}

void rg_launch_tprofile_svd(rg_plan *P, int nb, const double2 *gs1, int i0, int m, int j0, int nn, double2 *xwork,
                            double2 *vwork, double2 *w, double *tvals)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    hipLaunchKernelGGL(k_tprofile_svd, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->nmu, i0, m, j0, nn,
                       P->d_sqrtremu, xwork, vwork, w, tvals, P->d_status, 60);
    rg_timer_end(P, RGK_SVD, stop);
}

void rg_launch_colnorms(rg_plan *P, int nb, const double2 *y, int nv, double *out, size_t ostride)
{
    hipLaunchKernelGGL(k_colnorms, dim3(nb), dim3(256), 0, P->stream, y, P->np, P->nmu, nv, out, ostride);
}
