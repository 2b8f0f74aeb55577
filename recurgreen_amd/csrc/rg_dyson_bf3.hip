// rg_dyson_bf3.hip -- the Dyson product of the complex64 sweep on the BF16 matrix cores (gfx950).
//
// Gs1_new = -Gss diag(b) Gs1_old (src/waveguide.f90:438) with FP32 operands and FP32 accumulation,
// each real product carried out on v_mfma_f32_16x16x32_bf16 as the sum of six BF16 products.  An
// FP32 value x splits exactly into three BF16 values, x = h + m + l with h = bf16(x),
// m = bf16(x - h), l = bf16(x - h - m) (round to nearest even; the two differences are exact in
// FP32), so |m| <= 2^-8 |x| and |l| <= 2^-16 |x|.  x y = hh' + hm' + mh' + hl' + lh' + mm' + (the
// dropped ml' + lm' + ll', at most about 2^-23 |x y|, an FP32 rounding), and every BF16 product is
// exact in FP32; on random operands the result is as accurate as an FP32 product, but the
// instruction's summation cuts small products systematically (below).
//
// One workgroup of 8 wavefronts computes a realization's whole product, as the FP32 kernel:
// wavefront w owns row tiles {w/2, w/2 + 4} and column tiles {w%2 + 2y}.  k runs in chunks of 32
// through a double-buffered LDS stage of FP32 values ([row][k] for A, [col][k] for B, row length 36
// floats); each wavefront splits the fragments it reads.  147 KiB of LDS: one workgroup per CU, so
// the kernel is persistent (one workgroup per CU, realizations blockIdx.x + k gridDim.x): the
// chunk after the current one -- the next realization's first chunk after a realization's last --
// is loaded into registers while the current chunk is multiplied, which hides the loads at the
// start of each realization and lets its result stores drain under the next one's products.
// Measured at NP = 128 on 4096 random realizations: 0.50 ms per launch against 0.60-0.66 ms for the
// FP32 kernel; error against an FP64 product 1.5e-7 (max) and 1.6e-8 (rms) of sum |a||b| (FP32
// kernel: 2.1e-7 and 2.0e-8), no magnitude bias (9e-11).  The BF16 instruction issued about every
// 30 cycles here, so the products take about 0.31 ms of the launch.  In the N = 128 sweep over
// 1000 slices the mean transmission moved 4.7e-7 from the complex128 sweep's (FP32 kernel: 2.0e-8).
// The cause is the instruction's own summation: it aligns its 32 products to the largest and keeps
// one bit below FP32 precision, with no sticky bit, so the low bits of products much smaller than
// the largest are cut (two products 1 and 1.5 2^-24 sum to 1, where v_mfma_f32_16x16x4f32 gives
// 1 + 2^-23; through the C operand the same sum rounds correctly).  The rows of the sweep's
// matrices span many magnitudes, which makes the cut systematic; the kernel is therefore an option,
// not the default.
// MFMA lane layout (v_mfma_f32_16x16x32_bf16): A[i][8g + j] in lane i + 16g, element j;
// B[8g + j][c] in lane c + 16g, element j; result register v of lane l holds D[4(l/16) + v][l%16].
#include <algorithm>

#include "rg_internal.h"

namespace {

typedef float bf3_f4 __attribute__((ext_vector_type(4)));
typedef float bf3_f2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf3_b2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf3_b8 __attribute__((ext_vector_type(8)));
typedef unsigned bf3_u4 __attribute__((ext_vector_type(4)));

struct Bf3Pieces {
    bf3_b8 h, m, l;
};

template <int NP>
struct DysonSharedBf3 {
    static constexpr int S = 36;
    alignas(16) float a[2][2][NP][S];  // [buffer][re, im][row][k]
    alignas(16) float b[2][2][NP][S];  // [buffer][re, im][col][k]
};

__device__ __forceinline__ unsigned bf3_pack(float x, float y)
{
    const bf3_f2 v = {x, y};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf3_b2));
}

// x (8 consecutive k) -> h, m, l as BF16 x 8, element j in half j%2 of dword j/2
__device__ __forceinline__ Bf3Pieces bf3_split(const float4 &x0, const float4 &x1)
{
    const float x[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
    bf3_u4 h, m, l;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const unsigned hq = bf3_pack(x[2 * q], x[2 * q + 1]);
        const float r0 = x[2 * q] - __uint_as_float(hq << 16), r1 = x[2 * q + 1] - __uint_as_float(hq & 0xffff0000u);
        const unsigned mq = bf3_pack(r0, r1);
        const float s0 = r0 - __uint_as_float(mq << 16), s1 = r1 - __uint_as_float(mq & 0xffff0000u);
        h[q] = hq;
        m[q] = mq;
        l[q] = bf3_pack(s0, s1);
    }
    return {__builtin_bit_cast(bf3_b8, h), __builtin_bit_cast(bf3_b8, m), __builtin_bit_cast(bf3_b8, l)};
}

__device__ __forceinline__ Bf3Pieces bf3_neg(const Bf3Pieces &p)
{
    const bf3_u4 s = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    return {__builtin_bit_cast(bf3_b8, __builtin_bit_cast(bf3_u4, p.h) ^ s),
            __builtin_bit_cast(bf3_b8, __builtin_bit_cast(bf3_u4, p.m) ^ s),
            __builtin_bit_cast(bf3_b8, __builtin_bit_cast(bf3_u4, p.l) ^ s)};
}

// c += x y with the six BF16 products, smallest first
__device__ __forceinline__ bf3_f4 bf3_mac(const Bf3Pieces &x, const Bf3Pieces &y, bf3_f4 c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.m, y.m, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.l, y.h, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.h, y.l, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.m, y.h, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.h, y.m, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.h, y.h, c, 0, 0, 0);
    return c;
}

template <int NP>
__global__ __launch_bounds__(512, 2) void k_dyson_gemm_bf3(const float2 *__restrict__ gss, const float2 *__restrict__ gs1,
                                                           float2 *__restrict__ out, const double2 *__restrict__ bvec,
                                                           int n, int cs, int nb)
{
    static_assert(NP % 32 == 0, "k chunks of 32");
    constexpr int T = NP / 16, NK = NP / 32, NQ = NP * 16;  // tiles, k chunks, float4 per operand chunk
    constexpr int LOADS = NQ / 512;
    __shared__ DysonSharedBf3<NP> sh;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1, li = lane & 15, lg = lane >> 4;
    if ((int)blockIdx.x >= nb) return;  // the grid never exceeds the batch; uniform per workgroup

    // chunk kc of A: row idx >> 4, complex pair idx & 15 of the row's 32 k.  Of B: the 64 loads of
    // wave slot s = idx >> 6 cover k rows 8 (s & 3) + (lane & 7) and column pairs 8 (s >> 2) + (lane >> 3),
    // 128 contiguous bytes per k row, and land in distinct LDS banks when transposed.
    float4 ra[LOADS], rb[LOADS];
    auto load = [&](int r, int kc) __attribute__((always_inline)) {
        const float4 *A4 = reinterpret_cast<const float4 *>(gss + (size_t)r * NP * NP);
        const float4 *B4 = reinterpret_cast<const float4 *>(gs1 + (size_t)r * NP * NP);
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e, s = idx >> 6, l = idx & 63;
            ra[e] = A4[(size_t)(idx >> 4) * (NP / 2) + kc * 16 + (idx & 15)];
            rb[e] = B4[(size_t)(kc * 32 + 8 * (s & 3) + (l & 7)) * (NP / 2) + 8 * (s >> 2) + (l >> 3)];
        }
    };
    auto store = [&](int r, int kc, int buf) __attribute__((always_inline)) {
        const double2 *bv = bvec + (size_t)r * cs;  // cs = np: constants of each realization, else shared
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e, s = idx >> 6, l = idx & 63;
            const int row = idx >> 4, k2 = 2 * (idx & 15);
            *reinterpret_cast<float2 *>(&sh.a[buf][0][row][k2]) = make_float2(ra[e].x, ra[e].z);
            *reinterpret_cast<float2 *>(&sh.a[buf][1][row][k2]) = make_float2(ra[e].y, ra[e].w);
            const int k = 8 * (s & 3) + (l & 7), c = 2 * (8 * (s >> 2) + (l >> 3)), gk = kc * 32 + k;
            const double2 bk = gk < n ? bv[gk] : cmk(0.0, 0.0);
            sh.b[buf][0][c][k] = (float)(bk.x * rb[e].x - bk.y * rb[e].y);
            sh.b[buf][1][c][k] = (float)(bk.x * rb[e].y + bk.y * rb[e].x);
            sh.b[buf][0][c + 1][k] = (float)(bk.x * rb[e].z - bk.y * rb[e].w);
            sh.b[buf][1][c + 1][k] = (float)(bk.x * rb[e].w + bk.y * rb[e].z);
        }
    };
    auto frag = [&](const float (*plane)[DysonSharedBf3<NP>::S], int tile) __attribute__((always_inline)) {
        const float4 *p = reinterpret_cast<const float4 *>(&plane[16 * tile + li][8 * lg]);
        return bf3_split(p[0], p[1]);
    };

    bf3_f4 cr[2][4], ci[2][4];
    const int stride = gridDim.x;
    int buf = 0;
    load(blockIdx.x, 0);
    store(blockIdx.x, 0, 0);
    __syncthreads();
    for (int r = blockIdx.x; r < nb; r += stride) {
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) {
                cr[x][y] = bf3_f4{0.f, 0.f, 0.f, 0.f};
                ci[x][y] = bf3_f4{0.f, 0.f, 0.f, 0.f};
            }
        for (int kc = 0; kc < NK; kc++) {
            // the chunk after this one: (r, kc + 1), or the first of the next realization
            const int nr = kc + 1 < NK ? r : r + stride, nk = kc + 1 < NK ? kc + 1 : 0;
            const bool more = nr < nb;
            if (more) load(nr, nk);
            Bf3Pieces ar[2], ai[2], nai[2];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                const int rt = wr + 4 * x;
                if (rt >= T) continue;  // uniform per wavefront
                ar[x] = frag(sh.a[buf][0], rt);
                ai[x] = frag(sh.a[buf][1], rt);
                nai[x] = bf3_neg(ai[x]);
            }
#pragma unroll
            for (int y = 0; y < 4; y++) {  // one column tile at a time: fewer live fragments
                const int ct = wc + 2 * y;
                if (ct >= T) continue;  // uniform per wavefront
                const Bf3Pieces br = frag(sh.b[buf][0], ct), bi = frag(sh.b[buf][1], ct);
#pragma unroll
                for (int x = 0; x < 2; x++) {
                    if (wr + 4 * x >= T) continue;
                    cr[x][y] = bf3_mac(ar[x], br, cr[x][y]);
                    cr[x][y] = bf3_mac(nai[x], bi, cr[x][y]);
                    ci[x][y] = bf3_mac(ar[x], bi, ci[x][y]);
                    ci[x][y] = bf3_mac(ai[x], br, ci[x][y]);
                }
            }
            if (more) store(nr, nk, buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
        float2 *C = out + (size_t)r * NP * NP;
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
}

}  // namespace

bool rg_launch_gemm_bf3(rg_plan *P, int nb, const float2 *gss, const float2 *gs1_old, float2 *gs1_new)
{
    const int grid = std::max(1, std::min(nb, P->ncu));
    switch (P->np) {
        case 96:
            hipLaunchKernelGGL(k_dyson_gemm_bf3<96>, dim3(grid), dim3(512), 0, P->stream, gss, gs1_old, gs1_new, P->d_b,
                               P->nmu, P->cstride, nb);
            return true;
        case 128:
            hipLaunchKernelGGL(k_dyson_gemm_bf3<128>, dim3(grid), dim3(512), 0, P->stream, gss, gs1_old, gs1_new, P->d_b,
                               P->nmu, P->cstride, nb);
            return true;
        default: return false;
    }
}
