// rg_adjoint.hip -- kernels of the adjoint (reverse-mode) RGF sweep, complex128 (gfx950).
//
// The reference has no adjoint.  This differentiates its sweep (src/waveguide.f90:427-445),
// g_x = (A_x - B g_{x-1} B)^-1, S_x = -g_x B S_{x-1}, t = 2i s S_nx s (s = sqrt(mu)), for the losses
// L = Tr(t^H t) and L = T_max.  With adjoints defined by dL = Re Tr(Xbar^H dX):
//   Sbar_nx[i][j] = -4i (t W)_ab conj(s_i s_j), W = I or v v^H (v: top right singular vector of t),
//   x = nx..2:  gbar_x -= Sbar_x (B S_{x-1})^H;  Sbar_{x-1} = -B^H g_x^H Sbar_x;
//   x = 1:      gbar_1 += Sbar_1;
//   every x:    Cbar_x = -g_x^H gbar_x g_x^H;   gbar_{x-1} = -B^H Cbar_x B^H;
//   dL/du(x, y) = -Re sum_mn conj(Cbar_x[m][n]) conj(chi_m(y)) chi_n(y).
// The products run on rg_zgemm; this file holds the element-wise and projection kernels.  The
// host sequence (checkpoints every ~sqrt(nx) slices, segments recomputed forward) is in rg_plan.hip.
#include "rg_internal.h"

namespace {

// Loss L = Tr(t^H t) = sum |t_ab|^2 and its terminal adjoint Sbar_ij = -4i t_ab conj(s_i s_j).
template <class E>
__global__ __launch_bounds__(256) void k_adj_terminal(const E *__restrict__ gs1, int NP, int n, int imin, int imax,
                                                      int jmin, int jmax, const double2 *__restrict__ sqrtmu,
                                                      double *__restrict__ loss, E *__restrict__ sbar)
{
    __shared__ double s_red[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const E *G = gs1 + (size_t)b * NP * NP;
    E *SB = sbar + (size_t)b * NP * NP;
    double acc = 0.0;
    for (int e = tid; e < NP * NP; e += 256) {
        const int i = e / NP, j = e % NP;
        double2 out = cmk(0.0, 0.0);
        const int a = i - (imin - 1), c = j - (jmin - 1);
        if (i < n && j < n && a >= 0 && a < m && c >= 0 && c < nn) {
            const double2 sij = cmul(sqrtmu[i], sqrtmu[j]);
            const double2 v = cmul(sij, rg_mt<E>::ld(G[e]));
            const double2 t = cmk(-2.0 * v.y, 2.0 * v.x);  // 2i s_i G_ij s_j
            acc += t.x * t.x + t.y * t.y;
            const double2 q = cmul(t, cconj(sij));
            out = cmk(4.0 * q.y, -4.0 * q.x);  // -4i q
        }
        SB[e] = rg_mt<E>::make(out.x, out.y);
    }
    s_red[tid] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) s_red[tid] += s_red[tid + off];
        __syncthreads();
    }
    if (tid == 0) loss[b] = s_red[0];
}

// Loss T_max and its terminal adjoint from the top eigenvector x (unit norm) of the Gram matrix:
// gram_cols (A = t^H t, x = v): y = t v, T_max = |y|^2, t v v^H = y v^H;
// otherwise (A = t t^H, x = u): z = t^H u, T_max = |z|^2, u u^H t = u z^H.
template <class E>
__global__ __launch_bounds__(256) void k_adj_terminal_tmax(const E *__restrict__ gs1, int NP, int n, int imin,
                                                           int imax, int jmin, int jmax,
                                                           const double2 *__restrict__ sqrtmu,
                                                           const double2 *__restrict__ xvec, int xstride,
                                                           double *__restrict__ loss, E *__restrict__ sbar)
{
    __shared__ double2 s_x[256], s_y[256];
    __shared__ double s_red[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const bool gram_cols = nn <= m;
    const int nx_ = gram_cols ? nn : m, ny_ = gram_cols ? m : nn;
    const E *G = gs1 + (size_t)b * NP * NP;
    E *SB = sbar + (size_t)b * NP * NP;
    auto tval = [&](int a, int c) {  // t_ac = 2i s_i G_ij s_j
        const int i = imin - 1 + a, j = jmin - 1 + c;
        const double2 v = cmul(cmul(sqrtmu[i], rg_mt<E>::ld(G[(size_t)i * NP + j])), sqrtmu[j]);
        return cmk(-2.0 * v.y, 2.0 * v.x);
    };
    for (int i = tid; i < nx_; i += 256) s_x[i] = xvec[(size_t)b * xstride + i];
    __syncthreads();
    double acc = 0.0;
    for (int r = tid; r < ny_; r += 256) {
        double2 y = cmk(0.0, 0.0);
        if (gram_cols)
            for (int c = 0; c < nn; c++) y = cfma(y, tval(r, c), s_x[c]);
        else
            for (int a = 0; a < m; a++) y = cfma(y, cconj(tval(a, r)), s_x[a]);
        s_y[r] = y;
        acc += y.x * y.x + y.y * y.y;
    }
    s_red[tid] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) s_red[tid] += s_red[tid + off];
        __syncthreads();
    }
    if (tid == 0) loss[b] = s_red[0];
    for (int e = tid; e < NP * NP; e += 256) {
        const int i = e / NP, j = e % NP;
        double2 out = cmk(0.0, 0.0);
        const int a = i - (imin - 1), c = j - (jmin - 1);
        if (i < n && j < n && a >= 0 && a < m && c >= 0 && c < nn) {
            // tw_ac = y_a conj(v_c)  or  u_a conj(z_c)
            const double2 w = gram_cols ? cmul(s_y[a], cconj(s_x[c])) : cmul(s_x[a], cconj(s_y[c]));
            const double2 q = cmul(w, cconj(cmul(sqrtmu[i], sqrtmu[j])));
            out = cmk(4.0 * q.y, -4.0 * q.x);
        }
        SB[e] = rg_mt<E>::make(out.x, out.y);
    }
}

// gbar += sbar on the active n x n block.
template <class E>
__global__ __launch_bounds__(256) void k_adj_add(E *__restrict__ gbar, const E *__restrict__ sbar, int NP, int n)
{
    const size_t base = (size_t)blockIdx.x * NP * NP;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const size_t o = base + (size_t)(e / n) * NP + (e % n);
        const double2 r = cadd(rg_mt<E>::ld(gbar[o]), rg_mt<E>::ld(sbar[o]));
        gbar[o] = rg_mt<E>::make(r.x, r.y);
    }
}

// gbar = -conj(b)_m Cbar_mk conj(b)_k on the active block.
template <class E>
__global__ __launch_bounds__(256) void k_adj_scale(E *__restrict__ gbar, const E *__restrict__ cbar,
                                                   const double2 *__restrict__ bvec, int NP, int n)
{
    const size_t base = (size_t)blockIdx.x * NP * NP;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int m = e / n, k = e % n;
        const size_t o = base + (size_t)m * NP + k;
        const double2 r = cneg(cmul(cmul(cconj(bvec[m]), rg_mt<E>::ld(cbar[o])), cconj(bvec[k])));
        gbar[o] = rg_mt<E>::make(r.x, r.y);
    }
}

// dL/du(ixd, y) for one slice: e_q from Cbar, then grad(y) = -(1/ny) sum_q e_q tab[y][q].
//   Dirichlet (K_mn = (1/ny)[cos(|m-n| pi y) - cos((m+n+2) pi y)]):
//       e_q = sum_{|m-n|=q} Re Cbar_mn - sum_{m+n+2=q} Re Cbar_mn;
//   periodic (K_mn = (1/ny) exp(2 pi i (n-m) y), table cos/sin pairs):
//       e_2d = sum_{|n-m|=d} Re Cbar_mn,  e_2d+1 = sum_{n-m=d} Im Cbar_mn - sum_{m-n=d} Im Cbar_mn.
// grad has the potential layout (nxdiso, ny, r): grad[(b * ny + iy) * nxdiso + ixd].
template <class E>
__global__ __launch_bounds__(256) void k_adj_grad(const E *__restrict__ cbar, int NP, int n, int boundtype,
                                                  const double *__restrict__ tab, int ntp, int ntab, int ny, int nxdiso,
                                                  int ixd, double *__restrict__ grad)
{
    __shared__ double s_e[2 * 256 + 8];
    const int b = blockIdx.x, tid = threadIdx.x;
    const E *Cb = cbar + (size_t)b * NP * NP;
    for (int q = tid; q < ntab; q += 256) {
        double e = 0.0;
        if (boundtype == RG_DIRICHLET) {
            for (int m = 0; m < n; m++) {
                if (m + q < n) e += rg_mt<E>::ld(Cb[(size_t)m * NP + m + q]).x;           // n - m = q
                if (q > 0 && m - q >= 0) e += rg_mt<E>::ld(Cb[(size_t)m * NP + m - q]).x;  // m - n = q
                const int k = q - 2 - m;
                if (k >= 0 && k < n) e -= rg_mt<E>::ld(Cb[(size_t)m * NP + k]).x;
            }
        } else {
            const int d = q >> 1;
            for (int m = 0; m < n; m++) {
                if (q & 1) {
                    if (m + d < n) e += rg_mt<E>::ld(Cb[(size_t)m * NP + m + d]).y;
                    if (m - d >= 0) e -= rg_mt<E>::ld(Cb[(size_t)m * NP + m - d]).y;
                } else {
                    if (m + d < n) e += rg_mt<E>::ld(Cb[(size_t)m * NP + m + d]).x;
                    if (d > 0 && m - d >= 0) e += rg_mt<E>::ld(Cb[(size_t)m * NP + m - d]).x;
                }
            }
        }
        s_e[q] = e;
    }
    __syncthreads();
    const double sc = -1.0 / ny;
    for (int iy = tid; iy < ny; iy += 256) {
        const double *row = tab + (size_t)iy * ntp;
        double s = 0.0;
        for (int q = 0; q < ntab; q++) s = fma(s_e[q], row[q], s);
        grad[((size_t)b * ny + iy) * nxdiso + ixd] = sc * s;
    }
}

template <class E>
void adj_terminal(rg_plan *P, int nb, const E *gs1, double *loss, E *sbar)
{
    hipLaunchKernelGGL(k_adj_terminal<E>, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->nmu, P->imin, P->imax, P->jmin,
                       P->jmax, P->d_sqrtmu, loss, sbar);
}

template <class E>
void adj_terminal_tmax(rg_plan *P, int nb, const E *gs1, const double2 *xvec, int xstride, double *loss, E *sbar)
{
    hipLaunchKernelGGL(k_adj_terminal_tmax<E>, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->nmu, P->imin, P->imax,
                       P->jmin, P->jmax, P->d_sqrtmu, xvec, xstride, loss, sbar);
}

template <class E>
void adj_add(rg_plan *P, int nb, E *gbar, const E *sbar)
{
    hipLaunchKernelGGL(k_adj_add<E>, dim3(nb), dim3(256), 0, P->stream, gbar, sbar, P->np, P->nmu);
}

template <class E>
void adj_scale(rg_plan *P, int nb, E *gbar, const E *cbar)
{
    hipLaunchKernelGGL(k_adj_scale<E>, dim3(nb), dim3(256), 0, P->stream, gbar, cbar, P->d_b, P->np, P->nmu);
}

template <class E>
void adj_grad(rg_plan *P, int nb, const E *cbar, int ixd, double *grad)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_PROJECT, &stop);
    hipLaunchKernelGGL(k_adj_grad<E>, dim3(nb), dim3(256), 0, P->stream, cbar, P->np, P->nmu, P->p.boundtype, P->d_tab,
                       P->ntp, P->ntab, P->ny, P->nxdiso, ixd, grad);
    rg_timer_end(P, RGK_PROJECT, stop);
}

}  // namespace

void rg_launch_adj_terminal(rg_plan *P, int nb, const double2 *gs1, double *loss, double2 *sbar) { adj_terminal(P, nb, gs1, loss, sbar); }
void rg_launch_adj_terminal(rg_plan *P, int nb, const float2 *gs1, double *loss, float2 *sbar) { adj_terminal(P, nb, gs1, loss, sbar); }
void rg_launch_adj_terminal_tmax(rg_plan *P, int nb, const double2 *gs1, const double2 *xvec, int xstride, double *loss,
                                 double2 *sbar)
{
    adj_terminal_tmax(P, nb, gs1, xvec, xstride, loss, sbar);
}
void rg_launch_adj_terminal_tmax(rg_plan *P, int nb, const float2 *gs1, const double2 *xvec, int xstride, double *loss,
                                 float2 *sbar)
{
    adj_terminal_tmax(P, nb, gs1, xvec, xstride, loss, sbar);
}
void rg_launch_adj_add(rg_plan *P, int nb, double2 *gbar, const double2 *sbar) { adj_add(P, nb, gbar, sbar); }
void rg_launch_adj_add(rg_plan *P, int nb, float2 *gbar, const float2 *sbar) { adj_add(P, nb, gbar, sbar); }
void rg_launch_adj_scale(rg_plan *P, int nb, double2 *gbar, const double2 *cbar) { adj_scale(P, nb, gbar, cbar); }
void rg_launch_adj_scale(rg_plan *P, int nb, float2 *gbar, const float2 *cbar) { adj_scale(P, nb, gbar, cbar); }
void rg_launch_adj_grad(rg_plan *P, int nb, const double2 *cbar, int ixd, double *grad) { adj_grad(P, nb, cbar, ixd, grad); }
void rg_launch_adj_grad(rg_plan *P, int nb, const float2 *cbar, int ixd, double *grad) { adj_grad(P, nb, cbar, ixd, grad); }
