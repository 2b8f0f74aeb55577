// rg_dyson32.hip -- the Dyson product of the complex64 sweep, one realization per workgroup (gfx950).
//
// Gs1_new = -Gss diag(b) Gs1_old (src/waveguide.f90:438) for NP >= 80 on v_mfma_f32_16x16x4f32
// with FP32 accumulation.  One workgroup of 8 wavefronts computes the whole NP x NP product of one
// realization, so Gss and Gs1 are each read from HBM once (the 64 x 64 tiling read them twice).
// Wavefront w owns row tiles {w/2, w/2 + 4} and column tiles {w%2, w%2 + 2, w%2 + 4, w%2 + 6}
// (32 x 64 of the output for NP = 128).  k runs in chunks of 16: the next chunk is loaded into
// registers (16-byte global loads) while the current one is multiplied out of LDS, then stored to
// the other LDS buffer -- one barrier per chunk.  The B operand is scaled by b_k in FP64 before it
// is rounded to FP32, as in the 64 x 64 kernel.
//
// LDS planes are [k][row] (A, transposed on the way in) and [k][col] (B), row length NP + 4 so that
// the transposing stores of a wavefront (8 rows x 8 k-pairs) fall in distinct banks.
// MFMA lane layout (v_mfma_f32_16x16x4f32): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// result register v of lane l holds D[4(l/16) + v][l%16].
#include "rg_internal.h"

namespace {

typedef float dy_f4 __attribute__((ext_vector_type(4)));

template <int NP>
struct DysonShared {
    static constexpr int S = NP + 4;
    float a[2][2][16][S];  // [buffer][re, im][k][row]
    float b[2][2][16][S];  // [buffer][re, im][k][col]
};

template <int NP>
__global__ __launch_bounds__(512, 4) void k_dyson_gemm_wg(const float2 *__restrict__ gss, const float2 *__restrict__ gs1,
                                                          float2 *__restrict__ out, const double2 *__restrict__ bvec,
                                                          int n, int cs)
{
    bvec += (size_t)blockIdx.x * cs;  // cs = np: constants of each realization, else shared
    constexpr int T = NP / 16, NK = NP / 16, NQ = NP * 8;  // tiles, k chunks, float4 per operand chunk
    constexpr int LOADS = (NQ + 511) / 512;
    __shared__ DysonShared<NP> sh;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1, li = lane & 15, lg = lane >> 4;
    const size_t base = (size_t)blockIdx.x * NP * NP;
    const float4 *A4 = reinterpret_cast<const float4 *>(gss + base);
    const float4 *B4 = reinterpret_cast<const float4 *>(gs1 + base);

    float4 ra[LOADS], rb[LOADS];
    // chunk kc: A rows (idx >> 3), k pair (idx & 7); B k row idx / (NP/2), column pair idx % (NP/2)
    auto load = [&](int kc) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) {
                ra[e] = A4[(size_t)(idx >> 3) * (NP / 2) + kc * 8 + (idx & 7)];
                rb[e] = B4[(size_t)(kc * 16 + idx / (NP / 2)) * (NP / 2) + idx % (NP / 2)];
            }
        }
    };
    auto store = [&](int kc, int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) {
                const int row = idx >> 3, k2 = 2 * (idx & 7);
                sh.a[buf][0][k2][row] = ra[e].x;
                sh.a[buf][1][k2][row] = ra[e].y;
                sh.a[buf][0][k2 + 1][row] = ra[e].z;
                sh.a[buf][1][k2 + 1][row] = ra[e].w;
                const int kk = idx / (NP / 2), col = 2 * (idx % (NP / 2)), gk = kc * 16 + kk;
                const double2 bk = gk < n ? bvec[gk] : cmk(0.0, 0.0);
                const float2 v0 = make_float2((float)(bk.x * rb[e].x - bk.y * rb[e].y), (float)(bk.x * rb[e].y + bk.y * rb[e].x));
                const float2 v1 = make_float2((float)(bk.x * rb[e].z - bk.y * rb[e].w), (float)(bk.x * rb[e].w + bk.y * rb[e].z));
                *reinterpret_cast<float2 *>(&sh.b[buf][0][kk][col]) = make_float2(v0.x, v1.x);
                *reinterpret_cast<float2 *>(&sh.b[buf][1][kk][col]) = make_float2(v0.y, v1.y);
            }
        }
    };

    dy_f4 cr[2][4], ci[2][4];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            cr[x][y] = dy_f4{0.f, 0.f, 0.f, 0.f};
            ci[x][y] = dy_f4{0.f, 0.f, 0.f, 0.f};
        }

    load(0);
    store(0, 0);
    __syncthreads();
    for (int kc = 0; kc < NK; kc++) {
        const int buf = kc & 1;
        if (kc + 1 < NK) load(kc + 1);
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kr = 4 * ks + lg;
            float ar[2], ai[2], nai[2], br[4], bi[4];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                const int rt = wr + 4 * x;
                ar[x] = rt < T ? sh.a[buf][0][kr][16 * rt + li] : 0.f;
                ai[x] = rt < T ? sh.a[buf][1][kr][16 * rt + li] : 0.f;
                nai[x] = -ai[x];
            }
#pragma unroll
            for (int y = 0; y < 4; y++) {
                const int ct = wc + 2 * y;
                br[y] = ct < T ? sh.b[buf][0][kr][16 * ct + li] : 0.f;
                bi[y] = ct < T ? sh.b[buf][1][kr][16 * ct + li] : 0.f;
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) {
                    if (wr + 4 * x >= T || wc + 2 * y >= T) continue;  // uniform per wavefront
                    cr[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[x], br[y], cr[x][y], 0, 0, 0);
                    cr[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai[x], bi[y], cr[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[x], bi[y], ci[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[x], br[y], ci[x][y], 0, 0, 0);
                }
        }
        if (kc + 1 < NK) store(kc + 1, buf ^ 1);
        __syncthreads();
    }
    float2 *C = out + base;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int rt = wr + 4 * x, ct = wc + 2 * y;
            if (rt >= T || ct >= T) continue;
#pragma unroll
            for (int v = 0; v < 4; v++)
                C[(size_t)(16 * rt + 4 * lg + v) * NP + 16 * ct + li] = make_float2(-cr[x][y][v], -ci[x][y][v]);
        }
}

}  // namespace

bool rg_launch_gemm_wg(rg_plan *P, int nb, const float2 *gss, const float2 *gs1_old, float2 *gs1_new)
{
    switch (P->np) {
        case 80:
            hipLaunchKernelGGL(k_dyson_gemm_wg<80>, dim3(nb), dim3(512), 0, P->stream, gss, gs1_old, gs1_new, P->d_b, P->nmu, P->cstride);
            return true;
        case 96:
            hipLaunchKernelGGL(k_dyson_gemm_wg<96>, dim3(nb), dim3(512), 0, P->stream, gss, gs1_old, gs1_new, P->d_b, P->nmu, P->cstride);
            return true;
        case 112:
            hipLaunchKernelGGL(k_dyson_gemm_wg<112>, dim3(nb), dim3(512), 0, P->stream, gss, gs1_old, gs1_new, P->d_b, P->nmu, P->cstride);
            return true;
        case 128:
            hipLaunchKernelGGL(k_dyson_gemm_wg<128>, dim3(nb), dim3(512), 0, P->stream, gss, gs1_old, gs1_new, P->d_b, P->nmu, P->cstride);
            return true;
        default: return false;
    }
}
