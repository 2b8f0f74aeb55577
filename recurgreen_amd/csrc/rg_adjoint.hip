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

// Loss and terminal adjoint for L = Tr(t^H t) = sum |t_ab|^2 (kind 0), or for kind 1 with the
// weights tw = t v v^H supplied in `tw` (column-major per realization, m x nn, may be null for 0).
__global__ __launch_bounds__(256) void k_adj_terminal(const double2 *__restrict__ gs1, int NP, int n, int imin, int imax,
                                                      int jmin, int jmax, const double2 *__restrict__ sqrtmu,
                                                      const double2 *__restrict__ tw, double *__restrict__ loss,
                                                      double2 *__restrict__ sbar)
{
    __shared__ double s_red[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int m = imax - imin + 1, nn = jmax - jmin + 1;
    const double2 *G = gs1 + (size_t)b * NP * NP;
    double2 *SB = sbar + (size_t)b * NP * NP;
    double acc = 0.0;
    for (int e = tid; e < NP * NP; e += 256) {
        const int i = e / NP, j = e % NP;
        double2 out = cmk(0.0, 0.0);
        const int a = i - (imin - 1), c = j - (jmin - 1);
        if (i < n && j < n && a >= 0 && a < m && c >= 0 && c < nn) {
            const double2 sij = cmul(sqrtmu[i], sqrtmu[j]);
            const double2 v = cmul(sij, G[e]);
            const double2 t = cmk(-2.0 * v.y, 2.0 * v.x);  // 2i s_i G_ij s_j
            acc += t.x * t.x + t.y * t.y;
            const double2 w = tw ? tw[(size_t)b * m * nn + (size_t)c * m + a] : t;
            const double2 q = cmul(w, cconj(sij));
            out = cmk(4.0 * q.y, -4.0 * q.x);  // -4i q
        }
        SB[e] = out;
    }
    s_red[tid] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) s_red[tid] += s_red[tid + off];
        __syncthreads();
    }
    if (tid == 0 && !tw) loss[b] = s_red[0];
}

// gbar += sbar on the active n x n block.
__global__ __launch_bounds__(256) void k_adj_add(double2 *__restrict__ gbar, const double2 *__restrict__ sbar, int NP,
                                                 int n)
{
    const size_t base = (size_t)blockIdx.x * NP * NP;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const size_t o = base + (size_t)(e / n) * NP + (e % n);
        gbar[o] = cadd(gbar[o], sbar[o]);
    }
}

// gbar = -conj(b)_m Cbar_mk conj(b)_k on the active block.
__global__ __launch_bounds__(256) void k_adj_scale(double2 *__restrict__ gbar, const double2 *__restrict__ cbar,
                                                   const double2 *__restrict__ bvec, int NP, int n)
{
    const size_t base = (size_t)blockIdx.x * NP * NP;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int m = e / n, k = e % n;
        const size_t o = base + (size_t)m * NP + k;
        gbar[o] = cneg(cmul(cmul(cconj(bvec[m]), cbar[o]), cconj(bvec[k])));
    }
}

// dL/du(ixd, y) for one slice: e_q from Cbar, then grad(y) = -(1/ny) sum_q e_q tab[y][q].
//   Dirichlet (K_mn = (1/ny)[cos(|m-n| pi y) - cos((m+n+2) pi y)]):
//       e_q = sum_{|m-n|=q} Re Cbar_mn - sum_{m+n+2=q} Re Cbar_mn;
//   periodic (K_mn = (1/ny) exp(2 pi i (n-m) y), table cos/sin pairs):
//       e_2d = sum_{|n-m|=d} Re Cbar_mn,  e_2d+1 = sum_{n-m=d} Im Cbar_mn - sum_{m-n=d} Im Cbar_mn.
// grad has the potential layout (nxdiso, ny, r): grad[(b * ny + iy) * nxdiso + ixd].
__global__ __launch_bounds__(256) void k_adj_grad(const double2 *__restrict__ cbar, int NP, int n, int boundtype,
                                                  const double *__restrict__ tab, int ntp, int ntab, int ny, int nxdiso,
                                                  int ixd, double *__restrict__ grad)
{
    __shared__ double s_e[2 * 256 + 8];
    const int b = blockIdx.x, tid = threadIdx.x;
    const double2 *Cb = cbar + (size_t)b * NP * NP;
    for (int q = tid; q < ntab; q += 256) {
        double e = 0.0;
        if (boundtype == RG_DIRICHLET) {
            for (int m = 0; m < n; m++) {
                if (m + q < n) e += Cb[(size_t)m * NP + m + q].x;           // n - m = q
                if (q > 0 && m - q >= 0) e += Cb[(size_t)m * NP + m - q].x;  // m - n = q
                const int k = q - 2 - m;
                if (k >= 0 && k < n) e -= Cb[(size_t)m * NP + k].x;
            }
        } else {
            const int d = q >> 1;
            for (int m = 0; m < n; m++) {
                if (q & 1) {
                    if (m + d < n) e += Cb[(size_t)m * NP + m + d].y;
                    if (m - d >= 0) e -= Cb[(size_t)m * NP + m - d].y;
                } else {
                    if (m + d < n) e += Cb[(size_t)m * NP + m + d].x;
                    if (d > 0 && m - d >= 0) e += Cb[(size_t)m * NP + m - d].x;
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

}  // namespace

void rg_launch_adj_terminal(rg_plan *P, int nb, const double2 *gs1, const double2 *tw, double *loss, double2 *sbar)
{
    hipLaunchKernelGGL(k_adj_terminal, dim3(nb), dim3(256), 0, P->stream, gs1, P->np, P->nmu, P->imin, P->imax, P->jmin,
                       P->jmax, P->d_sqrtmu, tw, loss, sbar);
}

void rg_launch_adj_add(rg_plan *P, int nb, double2 *gbar, const double2 *sbar)
{
    hipLaunchKernelGGL(k_adj_add, dim3(nb), dim3(256), 0, P->stream, gbar, sbar, P->np, P->nmu);
}

void rg_launch_adj_scale(rg_plan *P, int nb, double2 *gbar, const double2 *cbar)
{
    hipLaunchKernelGGL(k_adj_scale, dim3(nb), dim3(256), 0, P->stream, gbar, cbar, P->d_b, P->np, P->nmu);
}

void rg_launch_adj_grad(rg_plan *P, int nb, const double2 *cbar, int ixd, double *grad)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_PROJECT, &stop);
    hipLaunchKernelGGL(k_adj_grad, dim3(nb), dim3(256), 0, P->stream, cbar, P->np, P->nmu, P->p.boundtype, P->d_tab,
                       P->ntp, P->ntab, P->ny, P->nxdiso, ixd, grad);
    rg_timer_end(P, RGK_PROJECT, stop);
}
