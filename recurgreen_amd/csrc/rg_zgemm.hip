// rg_zgemm.hip -- batched complex products with diagonal scalings on the matrix cores: complex128
// on v_mfma_f64_16x16x4 (rg_zgemm), complex64 on v_mfma_f32_16x16x4 (rg_cgemm32, the FP32 adjoint).
//
//   C = alpha * diag(dl) * op(A) * diag(dk) * op(B) * diag(dr) + beta * C,   op in {N, H}
//
// over a batch of np x np row-major matrices (np = 16 * T, the plan's padded order; the active
// order n <= np; rows and columns >= n are left untouched).  Used where the sweep needs products
// other than the hot Dyson update: the reflection Green functions of src/waveguide.f90:501-505 and
// every product of the adjoint sweep.  A workgroup of 4 wavefronts computes a 64x64 output tile
// (A[i][k] in lane i + 16k, B[k][j] in lane j + 16k; the result layout per element type is in
// rg_mt); k advances in chunks of 16 staged through LDS as separate re/im planes, with the
// conjugate transposes and the diagonal scalings applied (in FP64) while staging.
#include "rg_internal.h"

namespace {

template <int NP, class E>
__global__ __launch_bounds__(256) void k_zgemm(int n, double2 alpha, const E *__restrict__ A, int opA,
                                               const E *__restrict__ B, int opB, double2 beta,
                                               E *__restrict__ C, const double2 *__restrict__ dl,
                                               const double2 *__restrict__ dk, const double2 *__restrict__ dr)
{
    using R = typename rg_mt<E>::R;
    using zg_d4 = typename rg_mt<E>::V4;
    constexpr int TT = (NP + 63) / 64, LD = 64 + 2;
    __shared__ R AsR[16][LD], AsI[16][LD];  // op(A) chunk, [k][row]
    __shared__ R BsR[16][LD], BsI[16][LD];  // op(B) chunk, [k][col]
    const int b = blockIdx.y, ti0 = (blockIdx.x / TT) * 64, tj0 = (blockIdx.x % TT) * 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const E *Ab = A + (size_t)b * NP * NP;
    const E *Bb = B + (size_t)b * NP * NP;

    zg_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            cr[x][y] = zg_d4{0, 0, 0, 0};
            ci[x][y] = zg_d4{0, 0, 0, 0};
        }

    for (int k0 = 0; k0 < n; k0 += 16) {
        {   // op(A)[row][k]: thread -> row tid/4, k (tid%4)*4 .. +3
            const int rr = tid >> 2, kq = (tid & 3) * 4, gr = ti0 + rr;
            const double2 sl = (dl && gr < n) ? dl[gr] : cmk(1.0, 0.0);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gk = k0 + kq + e;
                double2 v = cmk(0.0, 0.0);
                if (gr < n && gk < n) {
                    v = opA ? cconj(rg_mt<E>::ld(Ab[(size_t)gk * NP + gr])) : rg_mt<E>::ld(Ab[(size_t)gr * NP + gk]);
                    if (dl) v = cmul(sl, v);
                }
                AsR[kq + e][rr] = v.x;
                AsI[kq + e][rr] = v.y;
            }
        }
        {   // op(B)[k][col]: thread -> k tid/16, cols (tid%16)*4 .. +3
            const int kk = tid >> 4, cq = (tid & 15) * 4, gk = k0 + kk;
            const double2 sk = (dk && gk < n) ? dk[gk] : cmk(1.0, 0.0);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gc = tj0 + cq + e;
                double2 v = cmk(0.0, 0.0);
                if (gk < n && gc < n) {
                    v = opB ? cconj(rg_mt<E>::ld(Bb[(size_t)gc * NP + gk])) : rg_mt<E>::ld(Bb[(size_t)gk * NP + gc]);
                    if (dk) v = cmul(sk, v);
                    if (dr) v = cmul(v, dr[gc]);
                }
                BsR[kk][cq + e] = v.x;
                BsI[kk][cq + e] = v.y;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kr = ks * 4 + lk;
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
    E *Cb = C + (size_t)b * NP * NP;
    const bool keep = beta.x != 0.0 || beta.y != 0.0;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int gr = ti0 + wr + 16 * x + rg_mt<E>::row(v, lk), gc = tj0 + wc + 16 * y + li;
                if (gr < n && gc < n) {
                    double2 r = cmul(alpha, cmk(cr[x][y][v], ci[x][y][v]));
                    if (keep) r = cfma(r, beta, rg_mt<E>::ld(Cb[(size_t)gr * NP + gc]));
                    Cb[(size_t)gr * NP + gc] = rg_mt<E>::make(r.x, r.y);
                }
            }
}

}  // namespace

template <class E>
static void zgemm_any(rg_plan *P, int nb, double2 alpha, const E *A, int opA, const E *B, int opB, double2 beta, E *C,
                      const double2 *dl, const double2 *dk, const double2 *dr)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_GEMM, &stop);
    const int tt = (P->np + 63) / 64;
    switch (P->np) {
#define RG_ZG(T)                                                                                             \
    case 16 * T:                                                                                             \
        hipLaunchKernelGGL((k_zgemm<16 * T, E>), dim3(tt * tt, nb), dim3(256), 0, P->stream, P->nmu, alpha, A, opA, B, \
                           opB, beta, C, dl, dk, dr);                                                        \
        break;
        RG_ZG(1)
        RG_ZG(2)
        RG_ZG(3)
        RG_ZG(4)
        RG_ZG(5)
        RG_ZG(6)
        RG_ZG(7)
        RG_ZG(8)
        RG_ZG(9)
        RG_ZG(10)
        RG_ZG(11)
        RG_ZG(12)
        RG_ZG(13)
        RG_ZG(14)
        RG_ZG(15)
        RG_ZG(16)
#undef RG_ZG
        default: break;
    }
    rg_timer_end(P, RGK_GEMM, stop);
}

void rg_zgemm(rg_plan *P, int nb, double2 alpha, const double2 *A, int opA, const double2 *B, int opB, double2 beta,
              double2 *C, const double2 *dl, const double2 *dk, const double2 *dr)
{
    zgemm_any<double2>(P, nb, alpha, A, opA, B, opB, beta, C, dl, dk, dr);
}

void rg_cgemm32(rg_plan *P, int nb, double2 alpha, const float2 *A, int opA, const float2 *B, int opB, double2 beta,
                float2 *C, const double2 *dl, const double2 *dk, const double2 *dr)
{
    zgemm_any<float2>(P, nb, alpha, A, opA, B, opB, beta, C, dl, dk, dr);
}
