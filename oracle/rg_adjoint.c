/*
 * rg_adjoint.c -- FP64 CPU adjoint of the RGF sweep (test infrastructure only; the checker the
 * GPU adjoint is compared against, itself checked against central finite differences of
 * rgo_transmission_matrix).
 *
 * The reference has no adjoint; this differentiates its sweep (src/waveguide.f90:427-445).
 * Forward, with B = diag(b) and the slice matrices A_x (src/waveguide.f90:349-384):
 *     g_1 = A_1^-1,  S_1 = g_1;
 *     g_x = (A_x - B g_{x-1} B)^-1,  S_x = -g_x B S_{x-1}      (x = 2..nx),
 *     t_ab = 2i s_i S_nx[i][j] s_j  (i = imin-1+a, j = jmin-1+b, s = sqrt(mu) complex).
 * Losses: L = Tr(t^H t) (kind 0) or L = T_max = sigma_max(t)^2 (kind 1).
 * With adjoints defined by dL = Re Tr(Xbar^H dX), the reverse sweep is
 *     Sbar_nx[i][j] = -4i (t W)_ab conj(s_i s_j),   W = I (kind 0) or v v^H (kind 1, v the top
 *                                                    right singular vector of t);
 *     for x = nx..2:  gbar_x += -Sbar_x (B S_{x-1})^H,   Sbar_{x-1} = -B^H g_x^H Sbar_x,
 *                     Cbar_x = -g_x^H gbar_x g_x^H,       gbar_{x-1} = -B^H Cbar_x B^H;
 *     x = 1:          gbar_1 += Sbar_1,  Cbar_1 = -g_1^H gbar_1 g_1^H,
 * where C_x = A_x - B g_{x-1} B.  A_x = A0_x - modal^H diag(u_x) modal, so
 *     dL/du(x, y) = -Re sum_mn conj(Cbar_x[m][n]) conj(chi_m(y)) chi_n(y).
 * upot and the gradient share the layout upot[iy * nxdiso + ix] (units U dx / k).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "rg_oracle.h"

/* C = alpha op(A) op(B) + beta C for n x n row-major; op: 0 = N, 1 = H. */
static void zgemm_nh(int n, zc alpha, const zc *A, int opA, const zc *B, int opB, zc beta, zc *C)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            zc s = 0;
            for (int k = 0; k < n; k++) {
                zc a = opA ? conj(A[(size_t)k * n + i]) : A[(size_t)i * n + k];
                zc b = opB ? conj(B[(size_t)j * n + k]) : B[(size_t)k * n + j];
                s += a * b;
            }
            C[(size_t)i * n + j] = alpha * s + (beta == 0 ? 0 : beta * C[(size_t)i * n + j]);
        }
}

/* Top eigenpair of the Hermitian n x n matrix H by shifted inverse iteration, the shift just above
 * the largest eigenvalue lam (from the singular values).  v has unit norm. */
static int top_eigvec(int n, const zc *H, double lam, zc *v)
{
    zc *M = (zc *)malloc(sizeof(zc) * (size_t)n * n), *work = (zc *)malloc(sizeof(zc) * (size_t)n * n);
    zc *x = (zc *)malloc(sizeof(zc) * n);
    int *ipiv = (int *)malloc(sizeof(int) * n);
    int st = (M && work && x && ipiv) ? RGO_OK : RGO_ENOMEM;
    if (!st) {
        double sigma = lam * (1.0 + 1e-10) + 1e-300;
        memcpy(M, H, sizeof(zc) * (size_t)n * n);
        for (int i = 0; i < n; i++) M[(size_t)i * n + i] -= sigma;
        st = rgo_inverse(n, M, ipiv, work);
        for (int i = 0; i < n; i++) v[i] = 1.0 / sqrt((double)n) * (1.0 + 0.01 * i);
        for (int it = 0; it < 6 && !st; it++) {
            double nrm = 0;
            for (int i = 0; i < n; i++) {
                zc s = 0;
                for (int k = 0; k < n; k++) s += M[(size_t)i * n + k] * v[k];
                x[i] = s;
                nrm += creal(s) * creal(s) + cimag(s) * cimag(s);
            }
            nrm = sqrt(nrm);
            for (int i = 0; i < n; i++) v[i] = x[i] / nrm;
        }
    }
    free(M); free(work); free(x); free(ipiv);
    return st;
}

int rgo_adjoint_grad(const rgo_wg *w, const double *upot, int loss_kind, double *val, double *grad)
{
    const int n = w->nmu, nx = w->nx;
    const size_t nn = (size_t)n * n;
    const int nr = w->imax - w->imin + 1, nc = w->jmax - w->jmin + 1;
    zc *g = (zc *)malloc(sizeof(zc) * nn * nx), *S = (zc *)malloc(sizeof(zc) * nn * nx);
    zc *A = (zc *)malloc(sizeof(zc) * nn), *work = (zc *)malloc(sizeof(zc) * nn);
    zc *T1 = (zc *)malloc(sizeof(zc) * nn), *T2 = (zc *)malloc(sizeof(zc) * nn);
    zc *gbar = (zc *)malloc(sizeof(zc) * nn), *Sbar = (zc *)malloc(sizeof(zc) * nn), *Cbar = (zc *)malloc(sizeof(zc) * nn);
    zc *t = (zc *)malloc(sizeof(zc) * (size_t)nr * nc), *tw = (zc *)malloc(sizeof(zc) * (size_t)nr * nc);
    int *ipiv = (int *)malloc(sizeof(int) * n);
    int st = (g && S && A && work && T1 && T2 && gbar && Sbar && Cbar && t && tw && ipiv) ? RGO_OK : RGO_ENOMEM;
    if (st) goto done;

    /* forward sweep, every slice kept */
    rgo_slice_matrix(w, 1, upot, g, work);
    if ((st = rgo_inverse(n, g, ipiv, work))) goto done;
    memcpy(S, g, sizeof(zc) * nn);
    for (int x = 2; x <= nx; x++) {
        zc *gx = g + nn * (x - 1), *gp = g + nn * (x - 2), *Sx = S + nn * (x - 1), *Sp = S + nn * (x - 2);
        rgo_slice_matrix(w, x, upot, A, work);
        for (int m = 0; m < n; m++)
            for (int k = 0; k < n; k++) gx[(size_t)m * n + k] = A[(size_t)m * n + k] - w->bmat[m] * gp[(size_t)m * n + k] * w->bmat[k];
        if ((st = rgo_inverse(n, gx, ipiv, work))) goto done;
        for (int m = 0; m < n; m++)
            for (int k = 0; k < n; k++) T1[(size_t)m * n + k] = w->bmat[m] * Sp[(size_t)m * n + k];
        zgemm_nh(n, -1.0, gx, 0, T1, 0, 0.0, Sx);
    }

    /* loss and terminal adjoint */
    const zc *GL = S + nn * (nx - 1);
    for (int a = 0; a < nr; a++)
        for (int b = 0; b < nc; b++) {
            int i = w->imin - 1 + a, j = w->jmin - 1 + b;
            t[(size_t)a * nc + b] = 2.0 * I * (csqrt(w->mumesh[i]) * GL[(size_t)i * n + j]) * csqrt(w->mumesh[j]);
        }
    if (loss_kind == 0) {
        double s2 = 0;
        for (size_t e = 0; e < (size_t)nr * nc; e++) s2 += creal(t[e]) * creal(t[e]) + cimag(t[e]) * cimag(t[e]);
        *val = s2;
        memcpy(tw, t, sizeof(zc) * (size_t)nr * nc);
    } else {
        /* H = t^H t (nc x nc); v = top eigenvector; tw = t v v^H */
        double *sv = (double *)malloc(sizeof(double) * (nr < nc ? nr : nc));
        zc *H = (zc *)malloc(sizeof(zc) * (size_t)nc * nc), *v = (zc *)malloc(sizeof(zc) * nc);
        st = (sv && H && v) ? rgo_singular_values(nr, nc, t, sv) : RGO_ENOMEM;
        if (!st) {
            for (int p = 0; p < nc; p++)
                for (int q = 0; q < nc; q++) {
                    zc s = 0;
                    for (int a = 0; a < nr; a++) s += conj(t[(size_t)a * nc + p]) * t[(size_t)a * nc + q];
                    H[(size_t)p * nc + q] = s;
                }
            *val = sv[0] * sv[0];
            st = top_eigvec(nc, H, sv[0] * sv[0], v);
        }
        if (!st)
            for (int a = 0; a < nr; a++) {
                zc tv = 0;
                for (int q = 0; q < nc; q++) tv += t[(size_t)a * nc + q] * v[q];
                for (int b = 0; b < nc; b++) tw[(size_t)a * nc + b] = tv * conj(v[b]);
            }
        free(sv); free(H); free(v);
        if (st) goto done;
    }
    memset(Sbar, 0, sizeof(zc) * nn);
    for (int a = 0; a < nr; a++)
        for (int b = 0; b < nc; b++) {
            int i = w->imin - 1 + a, j = w->jmin - 1 + b;
            Sbar[(size_t)i * n + j] = -4.0 * I * tw[(size_t)a * nc + b] * conj(csqrt(w->mumesh[i]) * csqrt(w->mumesh[j]));
        }
    memset(gbar, 0, sizeof(zc) * nn);
    memset(grad, 0, sizeof(double) * (size_t)w->nxdiso * w->ny);

    /* reverse sweep */
    for (int x = nx; x >= 1; x--) {
        zc *gx = g + nn * (x - 1);
        if (x >= 2) {
            zc *Sp = S + nn * (x - 2);
            for (int m = 0; m < n; m++)   /* T1 = B S_{x-1} */
                for (int k = 0; k < n; k++) T1[(size_t)m * n + k] = w->bmat[m] * Sp[(size_t)m * n + k];
            zgemm_nh(n, -1.0, Sbar, 0, T1, 1, 1.0, gbar);            /* gbar_x -= Sbar T1^H */
            zgemm_nh(n, -1.0, gx, 1, Sbar, 0, 0.0, T2);              /* T2 = -g_x^H Sbar */
            for (int m = 0; m < n; m++)
                for (int k = 0; k < n; k++) Sbar[(size_t)m * n + k] = conj(w->bmat[m]) * T2[(size_t)m * n + k];
        } else {
            for (size_t e = 0; e < nn; e++) gbar[e] += Sbar[e];
        }
        zgemm_nh(n, 1.0, gbar, 0, gx, 1, 0.0, T1);                   /* T1 = gbar g^H */
        zgemm_nh(n, -1.0, gx, 1, T1, 0, 0.0, Cbar);                  /* Cbar = -g^H gbar g^H */
        int id = x - w->nxfree;
        if (id >= 1 && id <= w->nxdiso)
            for (int iy = 0; iy < w->ny; iy++) {
                const zc *chi = w->modal + (size_t)iy * n;
                double s = 0;
                for (int m = 0; m < n; m++) {
                    zc r = 0;
                    for (int k = 0; k < n; k++) r += conj(Cbar[(size_t)m * n + k]) * chi[k];
                    s += creal(conj(chi[m]) * r);
                }
                grad[(size_t)iy * w->nxdiso + (id - 1)] = -s;
            }
        if (x >= 2)
            for (int m = 0; m < n; m++)
                for (int k = 0; k < n; k++)
                    gbar[(size_t)m * n + k] = -conj(w->bmat[m]) * Cbar[(size_t)m * n + k] * conj(w->bmat[k]);
    }
done:
    free(g); free(S); free(A); free(work); free(T1); free(T2); free(gbar); free(Sbar); free(Cbar);
    free(t); free(tw); free(ipiv);
    return st;
}
