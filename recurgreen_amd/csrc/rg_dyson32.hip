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

// ---- Three real products per complex product (option dyson = 3) ---------------------------------------
//
// The same product with the complex multiplication of Gauss: for A = Ar + i Ai, B = Br + i Bi,
//   P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi),  Re AB = P1 - P2,  Im AB = P3 - P1 - P2,
// three real products instead of four (25% fewer matrix-core instructions).  Each product is summed
// on its own FP32 accumulator; the sums Ar + Ai and Br + Bi are formed in FP32 from the staged
// operands.  The error of Im AB is bounded by eps32 (|P1| + |P2| + |P3|) sums, of the size of the
// four-product form's for operands of random phase.  The three accumulators of a wavefront's 2 x 4
// output tiles take 96 registers, so a workgroup of 8 wavefronts runs at two wavefronts per SIMD, one
// workgroup per CU: the kernel is persistent, and the chunk after the current one -- the next work
// item's first after an item's last -- is loaded into registers while the current one is multiplied,
// as in the BF16 kernel.  A work item is an MB x MB block of one realization's output (MB = NP up to
// 128; the four 128 x 128 quadrants at NP = 256, so that each operand element is read twice instead
// of four times by 64 x 64 tiles); items blockIdx.x + k gridDim.x.  k runs in chunks of 32 through a
// double-buffered LDS stage ([row][k] for A, [col][k] for B, row length 36 floats); k-step s of a
// chunk takes k = 8 g + s in lane group g, the same relabelling for both operands, so each lane reads
// its eight k values of a row (or column) as two 16-byte words.
template <int MB>
struct DysonShared3m {
    static constexpr int S = 36;
    alignas(16) float a[2][2][MB][S];  // [buffer][re, im][row][k]
    alignas(16) float b[2][2][MB][S];  // [buffer][re, im][col][k]
};

template <int NP, int MB>
__global__ __launch_bounds__(512, 2) void k_dyson_gemm_3m(const float2 *__restrict__ gss, const float2 *__restrict__ gs1,
                                                          float2 *__restrict__ out, const double2 *__restrict__ bvec,
                                                          int n, int cs, int nb)
{
    static_assert(NP % 32 == 0 && MB % 16 == 0 && NP % MB == 0 && MB <= 128, "k chunks of 32, 8 x 8 tiles at most");
    constexpr int T = MB / 16, NK = NP / 32, NQ = MB * 16;  // tiles, k chunks, float4 per operand chunk
    constexpr int QD = NP / MB, QI = QD * QD;               // blocks per dimension and per realization
    constexpr int LOADS = (NQ + 511) / 512;                 // rounds of 512 float4 (the last partial at MB 80, 112)
    __shared__ DysonShared3m<MB> sh;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1, li = lane & 15, lg = lane >> 4;
    const int items = nb * QI;
    if ((int)blockIdx.x >= items) return;  // the grid never exceeds the work items; uniform per workgroup

    // Item t: realization t / QI, output rows MB ((t % QI) / QD) .., columns MB ((t % QI) % QD) ...
    // Chunk kc of A: row idx >> 4 of the block, complex pair idx & 15 of the row's 32 k.  Of B: the 64
    // loads of wave slot s = idx >> 6 cover k rows 8 (s & 3) + (lane & 7) and column pairs
    // 8 (s >> 2) + (lane >> 3) of the block.  The next chunk's A and B are staged one after the other (A
    // loaded before the first half of the products and stored after it, then B), so that only one
    // operand's registers are in flight.
    float4 ra[LOADS], rb[LOADS];
    auto load_a = [&](int t, int kc) __attribute__((always_inline)) {
        const int r = t / QI, row0 = MB * ((t % QI) / QD);
        const float4 *A4 = reinterpret_cast<const float4 *>(gss + (size_t)r * NP * NP);
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (NQ % 512 != 0 && idx >= NQ) break;
            ra[e] = A4[(size_t)(row0 + (idx >> 4)) * (NP / 2) + kc * 16 + (idx & 15)];
        }
    };
    auto load_b = [&](int t, int kc) __attribute__((always_inline)) {
        const int r = t / QI, col0 = MB * ((t % QI) % QD);
        const float4 *B4 = reinterpret_cast<const float4 *>(gs1 + (size_t)r * NP * NP);
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e, s = idx >> 6, l = idx & 63;
            if (NQ % 512 != 0 && idx >= NQ) break;
            rb[e] = B4[(size_t)(kc * 32 + 8 * (s & 3) + (l & 7)) * (NP / 2) + col0 / 2 + 8 * (s >> 2) + (l >> 3)];
        }
    };
    auto store_a = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (NQ % 512 != 0 && idx >= NQ) break;
            const int row = idx >> 4, k2 = 2 * (idx & 15);
            *reinterpret_cast<float2 *>(&sh.a[buf][0][row][k2]) = make_float2(ra[e].x, ra[e].z);
            *reinterpret_cast<float2 *>(&sh.a[buf][1][row][k2]) = make_float2(ra[e].y, ra[e].w);
        }
    };
    auto store_b = [&](int t, int kc, int buf) __attribute__((always_inline)) {
        const double2 *bv = bvec + (size_t)(t / QI) * cs;  // cs = np: constants of each realization, else shared
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e, s = idx >> 6, l = idx & 63;
            if (NQ % 512 != 0 && idx >= NQ) break;
            const int k = 8 * (s & 3) + (l & 7), c = 2 * (8 * (s >> 2) + (l >> 3)), gk = kc * 32 + k;
            const double2 bk = gk < n ? bv[gk] : cmk(0.0, 0.0);
            sh.b[buf][0][c][k] = (float)(bk.x * rb[e].x - bk.y * rb[e].y);
            sh.b[buf][1][c][k] = (float)(bk.x * rb[e].y + bk.y * rb[e].x);
            sh.b[buf][0][c + 1][k] = (float)(bk.x * rb[e].z - bk.y * rb[e].w);
            sh.b[buf][1][c + 1][k] = (float)(bk.x * rb[e].w + bk.y * rb[e].z);
        }
    };
    // the eight k values 8 lg .. 8 lg + 7 of row (or column) 16 tile + li of a plane
    auto frag = [&](const float (*plane)[DysonShared3m<MB>::S], int tile, float (&v)[8]) __attribute__((always_inline)) {
        const float4 *p = reinterpret_cast<const float4 *>(&plane[16 * tile + li][8 * lg]);
        const float4 x0 = p[0], x1 = p[1];
        v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w;
        v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
    };

    dy_f4 p1[2][4], p2[2][4], p3[2][4];
    const int stride = gridDim.x;
    int buf = 0;
    load_a(blockIdx.x, 0);
    load_b(blockIdx.x, 0);
    store_a(0);
    store_b(blockIdx.x, 0, 0);
    __syncthreads();
    for (int t = blockIdx.x; t < items; t += stride) {
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) {
                p1[x][y] = dy_f4{0.f, 0.f, 0.f, 0.f};
                p2[x][y] = dy_f4{0.f, 0.f, 0.f, 0.f};
                p3[x][y] = dy_f4{0.f, 0.f, 0.f, 0.f};
            }
#pragma nounroll
        for (int kc = 0; kc < NK; kc++) {
            // the chunk after this one: (t, kc + 1), or the first of the next item
            const int nt = kc + 1 < NK ? t : t + stride, nk = kc + 1 < NK ? kc + 1 : 0;
            const bool more = nt < items;
            if (more) load_a(nt, nk);
            float ar[2][8], ai[2][8];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                const int rt = wr + 4 * x;
                if (rt >= T) continue;  // uniform per wavefront
                frag(sh.a[buf][0], rt, ar[x]);
                frag(sh.a[buf][1], rt, ai[x]);
            }
#pragma unroll
            for (int y = 0; y < 4; y++) {  // one column tile at a time: fewer live fragments
                const int ct = wc + 2 * y;
                if (ct >= T) continue;  // uniform per wavefront
                float br[8], bi[8];
                frag(sh.b[buf][0], ct, br);
                frag(sh.b[buf][1], ct, bi);
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const float bs = br[j] + bi[j];  // the sums are formed where they are used
#pragma unroll
                    for (int x = 0; x < 2; x++) {
                        if (wr + 4 * x >= T) continue;
                        p1[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[x][j], br[j], p1[x][y], 0, 0, 0);
                        p2[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[x][j], bi[j], p2[x][y], 0, 0, 0);
                        p3[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[x][j] + ai[x][j], bs, p3[x][y], 0, 0, 0);
                    }
                }
                // keeps the next column tile's fragment reads below this one's products (register pressure)
                __builtin_amdgcn_sched_barrier(0);
                if (y == 1 && more) {  // the other buffer was last read before the previous barrier
                    store_a(buf ^ 1);
                    load_b(nt, nk);
                }
            }
            if (more) store_b(nt, nk, buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
        const int r = t / QI, row0 = MB * ((t % QI) / QD), col0 = MB * ((t % QI) % QD);
        float2 *C = out + (size_t)r * NP * NP;
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) {
                const int rt = wr + 4 * x, ct = wc + 2 * y;
                if (rt >= T || ct >= T) continue;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float re = p1[x][y][v] - p2[x][y][v], im = p3[x][y][v] - p1[x][y][v] - p2[x][y][v];
                    C[(size_t)(row0 + 16 * rt + 4 * lg + v) * NP + col0 + 16 * ct + li] = make_float2(-re, -im);
                }
            }
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

// Option dyson = 3: NP 96, 128 (the whole matrix per work item), 160, 192, 224 and 256 (four blocks of 80 to 128
// per realization); false otherwise.  Blocks of 80 and 112 rows (5 x 5 and 7 x 7 tiles over the 8 waves' 4 x 2
// grid of up to 8 x 8) stage their operands in a last, partial round of 512 float4.  At NP 80 and 112 themselves
// (k chunks of 32 would need zero padding) the four-product kernel with one workgroup per realization is
// faster (measured with padding):
// 0.42 against 0.55 ms and 0.96 against 1.16 ms per launch of 8192; the blocks of 80 and 112 beat the 64 x 64
// tiles at NP 160 and 224 (1.52 against 1.89 ms and 3.21 against 4.50 ms per launch of 4096).
bool rg_launch_gemm_3m(rg_plan *P, int nb, const float2 *gss, const float2 *gs1_old, float2 *gs1_new)
{
    const int qi = P->np > 128 ? 4 : 1, items = nb * qi;  // the orders above 128 as 2 x 2 blocks
    const int grid = items < P->ncu ? (items < 1 ? 1 : items) : P->ncu;
    switch (P->np) {
#define RG_3M(N, MB)                                                                                           \
    case N:                                                                                                    \
        hipLaunchKernelGGL((k_dyson_gemm_3m<N, MB>), dim3(grid), dim3(512), 0, P->stream, gss, gs1_old, gs1_new, \
                           P->d_b, P->nmu, P->cstride, nb);                                                    \
        return true;
        RG_3M(96, 96)
        RG_3M(128, 128)
        RG_3M(160, 80)
        RG_3M(192, 96)
        RG_3M(224, 112)
        RG_3M(256, 128)
#undef RG_3M
        default: return false;
    }
}
