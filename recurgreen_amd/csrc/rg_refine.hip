// rg_refine.hip -- FP64 iterative refinement of G(L;0) for the complex64 sweeps (plan option refine = k).
//
// G(L;0) is the last block of X = H^-1 E_1 for the block-tridiagonal H of the guide: diagonal blocks
// H_xx = diag(a_x) - V_x (a_x the slice's diagonal, first, bulk or last; V_x the projected disorder) and
// off-diagonal blocks diag(b).  The sweep of src/waveguide.f90:424-440, Gss_x = (H_xx - B Gss_{x-1} B)^-1 and
// Gs1_x = -Gss_x B Gs1_{x-1}, is the forward elimination of H X = E_1 with the inverse Schur complements
// g_x = Gss_x.  The complex64 sweep keeps every slice's g_x; then, from X = 0, k + 1 passes of
//   R = E_1 - H X                                   (FP64; V_x X_x on the FP64 matrix cores; R = E_1 at X = 0)
//   y_1 = g_1 R_1,   y_x = g_x (R_x - b y_{x-1})       (forward, x = 2..nx)
//   d_nx = y_nx,     d_x = y_x - g_x b d_{x+1}         (backward)
//   X <- X + d
// with the complex64 g_x taken into FP64 products.  The first pass is the complex64 sweep's own solution
// (its forward half, y = Gs1 at every slice, kept from the sweep);
// each further pass multiplies the error by about eps32 times the growth of the elimination, so X, and
// G(L;0) = X_nx, reach the accuracy of an FP64 sweep of the same H in one or two passes.  The FP64
// Fisher-Lee relation and eigenvalues follow (rg_plan.hip).
//
// Storage, per realization, of a pass: g, X and R for every slice, complex128 np x np each
// ([slice][realization][np][np]), and the slice coefficients ([slice][realization][ntab]); the pass size
// (ref_cap) is set from the free device memory on first use.
#include <algorithm>

#include "rg_internal.h"

namespace {

constexpr int kEw = 256;  // threads of the element-wise kernels

int ew_blocks(size_t elems) { return (int)std::min<size_t>(64, (elems + kEw - 1) / kEw); }

// g_x: the sweep's complex64 inverse to complex128; y_x = Gs1_x, the sweep's forward solution of
// H X = E_1 (the first pass's forward half), to complex128 on the n x n block (zero padding); the
// slice's summed coefficients (zeros on slices without disorder).  Grid (blocks over np^2, realizations).
__global__ __launch_bounds__(kEw) void k_ref_store(const float2 *__restrict__ g32, double2 *__restrict__ g,
                                                   const float2 *__restrict__ y32, double2 *__restrict__ y,
                                                   const rg_cqv cqv, double *__restrict__ cq, int n, int np, int ntab)
{
    const int b = blockIdx.y, tid = threadIdx.x, np2 = np * np;
    const size_t o = (size_t)b * np2;
    for (int e = blockIdx.x * kEw + tid; e < np2; e += gridDim.x * kEw) {
        const float2 v = g32[o + e], w = y32[o + e];
        g[o + e] = cmk((double)v.x, (double)v.y);
        y[o + e] = (e / np < n && e % np < n) ? cmk((double)w.x, (double)w.y) : cmk(0.0, 0.0);
    }
    if (blockIdx.x == 0) {
        double *dst = cq + (size_t)b * ntab;
        if (cqv.cq) rg_load_cq(dst, cqv, b, tid, kEw);
        else
            for (int q = tid; q < ntab; q += kEw) dst[q] = 0.0;
    }
}

// V_x of item y = (slice, realization) from its coefficients, as the slice inversions form it
// (Dirichlet: real symmetric Toeplitz minus Hankel; periodic: Hermitian Toeplitz).  Grid (blocks, items).
__global__ __launch_bounds__(kEw) void k_ref_vmat(const double *__restrict__ cqs, int ntab, int boundtype, int n, int np,
                                                  double2 *__restrict__ V)
{
    const size_t item = blockIdx.y;
    const double *c = cqs + item * ntab;
    double2 *Vm = V + item * np * np;
    for (int e = blockIdx.x * kEw + threadIdx.x; e < np * np; e += gridDim.x * kEw) {
        const int r = e / np, col = e % np;
        double2 v = cmk(0.0, 0.0);
        if (r < n && col < n) {
            if (boundtype == RG_DIRICHLET) {
                v.x = c[abs(r - col)] - c[r + col + 2];
            } else {
                const int d = col - r;
                v = d >= 0 ? cmk(c[2 * d], c[2 * d + 1]) : cmk(c[-2 * d], -c[-2 * d + 1]);
            }
        }
        Vm[e] = v;
    }
}

// R_x (holding V_x X_x) <- E_1 - H X on slice x: + I on the first slice, - diag(a_x) X_x, - diag(b)
// (X_{x-1} + X_{x+1}).  Items y = (slice x0 + y / nb, realization y % nb); grid (blocks, items).
__global__ __launch_bounds__(kEw) void k_ref_resid(double2 *__restrict__ R, const double2 *__restrict__ X,
                                                   const double2 *__restrict__ adiag, const double2 *__restrict__ bvec,
                                                   int cs, int nb, int x0, int nx, int n, int np)
{
    const int item = blockIdx.y, b = item % nb, x = x0 + item / nb;
    const size_t np2 = (size_t)np * np, slab = (size_t)nb * np2, o = (size_t)x * slab + (size_t)b * np2;
    const int slot = x == 0 ? 0 : (x == nx - 1 ? 2 : 1);
    const double2 *a = adiag + (size_t)b * 3 * cs + (size_t)slot * np;
    const double2 *bv = bvec + (size_t)b * cs;
    for (int e = blockIdx.x * kEw + threadIdx.x; e < np * np; e += gridDim.x * kEw) {
        const int r = e / np, col = e % np;
        if (r >= n || col >= n) continue;
        double2 nbr = cmk(0.0, 0.0);
        if (x > 0) nbr = cadd(nbr, X[o - slab + e]);
        if (x < nx - 1) nbr = cadd(nbr, X[o + slab + e]);
        double2 v = csub(csub(R[o + e], cmul(a[r], X[o + e])), cmul(bv[r], nbr));
        if (x == 0 && r == col) v.x += 1.0;
        R[o + e] = v;
    }
}

// W = Pm - diag(b) Q per realization (Pm or Q null: zero); with Xq, also Xq += Q (the backward half adds
// each finished d_{x+1} to X as it reads it).  Grid (blocks, realizations).
__global__ __launch_bounds__(kEw) void k_ref_rhs(const double2 *__restrict__ Pm, const double2 *__restrict__ Q,
                                                 const double2 *__restrict__ bvec, int cs, double2 *__restrict__ W, int n,
                                                 int np, double2 *__restrict__ Xq)
{
    const int b = blockIdx.y;
    const size_t o = (size_t)b * np * np;
    const double2 *bv = bvec + (size_t)b * cs;
    for (int e = blockIdx.x * kEw + threadIdx.x; e < np * np; e += gridDim.x * kEw) {
        const int r = e / np, col = e % np;
        if (r >= n || col >= n) continue;
        double2 v = Pm ? Pm[o + e] : cmk(0.0, 0.0);
        if (Q) {
            const double2 q = Q[o + e];
            v = csub(v, cmul(bv[r], q));
            if (Xq) Xq[o + e] = cadd(Xq[o + e], q);
        }
        W[o + e] = v;
    }
}

__global__ __launch_bounds__(kEw) void k_ref_add(double2 *__restrict__ X, const double2 *__restrict__ D, size_t count)
{
    for (size_t e = (size_t)blockIdx.x * kEw + threadIdx.x; e < count; e += (size_t)gridDim.x * kEw) X[e] = cadd(X[e], D[e]);
}

}  // namespace

rg_status rg_refine_reserve(rg_plan *P)
{
    if (P->d_ref_g) return RG_OK;
    const size_t np2 = (size_t)P->np * P->np, nx = (size_t)P->nx, mat = np2 * sizeof(double2);
    size_t fr = 0, total = 0;
    if (hipMemGetInfo(&fr, &total) != hipSuccess) return RG_EDEVICE;
    // V blocks of one chunk of slices: about 1 GiB, at most 65535 (one grid dimension)
    const size_t vcap0 = std::max<size_t>(1, std::min<size_t>(65535, ((size_t)1 << 30) / mat));
    const size_t per = 3 * nx * mat + nx * (size_t)P->ntab * sizeof(double) + 2 * mat;  // g, X, R, cq, W, V
    const size_t budget = fr / 10 * 6;
    if (budget <= vcap0 * mat + per) return RG_ENOMEM;
    size_t cap = std::min<size_t>({(size_t)P->batch_max, 65535, (budget - vcap0 * mat) / per});
    if (P->refine_cap > 0) cap = std::min<size_t>(cap, (size_t)P->refine_cap);
    const size_t vcap = std::max(vcap0, cap);
    auto alloc = [](auto **p, size_t count) {
        return hipMalloc((void **)p, count * sizeof(**p)) == hipSuccess;
    };
    if (!alloc(&P->d_ref_g, cap * nx * np2) || !alloc(&P->d_ref_x, cap * nx * np2) || !alloc(&P->d_ref_r, cap * nx * np2) ||
        !alloc(&P->d_ref_cq, cap * nx * P->ntab) || !alloc(&P->d_ref_w, cap * np2) || !alloc(&P->d_ref_v, vcap * np2)) {
        (void)hipGetLastError();
        return RG_ENOMEM;
    }
    // padding rows and columns are never written by the passes and stay zero
    if (hipMemset(P->d_ref_r, 0, cap * nx * mat) != hipSuccess || hipMemset(P->d_ref_w, 0, cap * mat) != hipSuccess)
        return RG_EDEVICE;
    P->ref_cap = (int)cap;
    P->ref_vcap = (int)vcap;
    return RG_OK;
}

void rg_refine_release(rg_plan *P)
{
    for (void *p : {(void *)P->d_ref_g, (void *)P->d_ref_x, (void *)P->d_ref_r, (void *)P->d_ref_w, (void *)P->d_ref_v,
                    (void *)P->d_ref_cq})
        if (p) (void)hipFree(p);
    P->d_ref_g = P->d_ref_x = P->d_ref_r = P->d_ref_w = P->d_ref_v = nullptr;
    P->d_ref_cq = nullptr;
    P->ref_cap = P->ref_vcap = 0;
}

void rg_refine_store(rg_plan *P, int nb, int ix, bool with_disorder, const float2 *gs1)
{
    const size_t np2 = (size_t)P->np * P->np, x = (size_t)ix - 1;
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    hipLaunchKernelGGL(k_ref_store, dim3(ew_blocks(np2), nb), dim3(kEw), 0, P->stream, P->d_gss32,
                       P->d_ref_g + x * nb * np2, gs1, P->d_ref_r + x * nb * np2, cq, P->d_ref_cq + x * nb * P->ntab,
                       P->nmu, P->np, P->ntab);
}

const double2 *rg_refine_solve(rg_plan *P, int nb)
{
    const int n = P->nmu, np = P->np, nx = P->nx;
    const size_t np2 = (size_t)np * np, slab = (size_t)nb * np2, all = (size_t)nx * slab;
    const double2 one = cmk(1.0, 0.0), zero = cmk(0.0, 0.0);
    double2 *X = P->d_ref_x, *R = P->d_ref_r, *W = P->d_ref_w;
    const double2 *G = P->d_ref_g;
    const int gx = ew_blocks(np2), chunk = std::max(1, P->ref_vcap / nb);
    (void)hipMemsetAsync(X, 0, all * sizeof(double2), P->stream);
    // pass 0: R holds the sweep's forward solution y (stored slice by slice); X = 0
    for (int pass = 0; pass <= P->refine; pass++) {
        if (pass > 0) {
            // R = E_1 - H X
            for (int x0 = 0; x0 < nx; x0 += chunk) {
                const int cnt = std::min(chunk, nx - x0), items = cnt * nb;
                hipLaunchKernelGGL(k_ref_vmat, dim3(gx, items), dim3(kEw), 0, P->stream,
                                   P->d_ref_cq + (size_t)x0 * nb * P->ntab, P->ntab, P->p.boundtype, n, np, P->d_ref_v);
                rg_zgemm(P, items, n, n, n, one, P->d_ref_v, np, np2, X + x0 * slab, np, np2, zero, R + x0 * slab, np, np2);
                hipLaunchKernelGGL(k_ref_resid, dim3(gx, items), dim3(kEw), 0, P->stream, R, X, P->d_adiag, P->d_b,
                                   P->cstride, nb, x0, nx, n, np);
            }
        }
        // forward: R_x <- y_x
        for (int x = 0; x < nx && pass > 0; x++) {
            hipLaunchKernelGGL(k_ref_rhs, dim3(gx, nb), dim3(kEw), 0, P->stream, R + x * slab,
                               x > 0 ? R + (x - 1) * slab : nullptr, P->d_b, P->cstride, W, n, np, nullptr);
            rg_zgemm(P, nb, n, n, n, one, G + x * slab, np, np2, W, np, np2, zero, R + x * slab, np, np2);
        }
        // backward: R_x <- d_x = y_x + g_x (-b d_{x+1}), and X_{x+1} += d_{x+1}; then X_1 += d_1
        for (int x = nx - 2; x >= 0; x--) {
            hipLaunchKernelGGL(k_ref_rhs, dim3(gx, nb), dim3(kEw), 0, P->stream, nullptr, R + (x + 1) * slab, P->d_b,
                               P->cstride, W, n, np, X + (x + 1) * slab);
            rg_zgemm(P, nb, n, n, n, one, G + x * slab, np, np2, W, np, np2, one, R + x * slab, np, np2);
        }
        hipLaunchKernelGGL(k_ref_add, dim3((unsigned)std::min<size_t>(4096, (slab + kEw - 1) / kEw)), dim3(kEw), 0,
                           P->stream, X, (const double2 *)R, slab);
    }
    return X + (size_t)(nx - 1) * slab;
}
