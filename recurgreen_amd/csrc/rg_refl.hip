// rg_refl.hip -- Fisher-Lee relations of the reflection/transmission sweep (gfx950).
//
// src/waveguide.f90:508-512: with s_n = sqrt(Re mu_n) (real: the real part drops evanescent
// modes), t = 2i s Gs1 s and r = -I + 2i s G11 s, full nmu x nmu, written column-major per
// realization as the C ABI lays matrices out.
#include "rg_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_refl_trans_out(const double2 *__restrict__ g11, const double2 *__restrict__ gs1,
                                                        int NP, int n, const double *__restrict__ sre,
                                                        double2 *__restrict__ rmat, double2 *__restrict__ tmat)
{
    const int b = blockIdx.x;
    const double2 *G11 = g11 + (size_t)b * NP * NP, *GS1 = gs1 + (size_t)b * NP * NP;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e % n, j = e / n;
        const double s = sre[i] * sre[j];
        const double2 gt = GS1[(size_t)i * NP + j], gr = G11[(size_t)i * NP + j];
        tmat[(size_t)b * n * n + e] = cmk(-2.0 * s * gt.y, 2.0 * s * gt.x);
        rmat[(size_t)b * n * n + e] = cmk(-2.0 * s * gr.y - (i == j ? 1.0 : 0.0), 2.0 * s * gr.x);
    }
}

// Deposition matrix (src/waveguide.f90:741-742): dep = 2i s G s for one block G(fpos; 1), full
// nmu x nmu, column-major per realization.
__global__ __launch_bounds__(256) void k_deposition_out(const double2 *__restrict__ g, int NP, int n,
                                                        const double *__restrict__ sre, double2 *__restrict__ dep)
{
    const int b = blockIdx.x;
    const double2 *G = g + (size_t)b * NP * NP;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e % n, j = e / n;
        const double s = sre[i] * sre[j];
        const double2 gv = G[(size_t)i * NP + j];
        dep[(size_t)b * n * n + e] = cmk(-2.0 * s * gv.y, 2.0 * s * gv.x);
    }
}

// a += b over count complex entries.
__global__ __launch_bounds__(256) void k_add_c128(double2 *__restrict__ a, const double2 *__restrict__ b, size_t count)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        a[e] = cadd(a[e], b[e]);
}

}  // namespace

void rg_launch_deposition_out(rg_plan *P, int nb, const double2 *g, double2 *dep)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    hipLaunchKernelGGL(k_deposition_out, dim3(nb), dim3(256), 0, P->stream, g, P->np, P->nmu, P->d_sqrtremu, dep);
    rg_timer_end(P, RGK_SVD, stop);
}

void rg_launch_add_c128(rg_plan *P, double2 *a, const double2 *b, size_t count)
{
    const size_t blocks = std::min<size_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL(k_add_c128, dim3((unsigned)blocks), dim3(256), 0, P->stream, a, b, count);
}

void rg_launch_refl_trans_out(rg_plan *P, int nb, const double2 *g11, const double2 *gs1, double2 *rmat, double2 *tmat)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_SVD, &stop);
    hipLaunchKernelGGL(k_refl_trans_out, dim3(nb), dim3(256), 0, P->stream, g11, gs1, P->np, P->nmu, P->d_sqrtremu,
                       rmat, tmat);
    rg_timer_end(P, RGK_SVD, stop);
}
