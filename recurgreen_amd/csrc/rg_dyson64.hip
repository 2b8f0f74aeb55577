// rg_dyson64.hip -- the Dyson product of the complex128 sweep, one realization per workgroup (gfx950).
//
// Gs1_new = -Gss diag(b) Gs1_old (src/waveguide.f90:438) for NP = 80..128 on v_mfma_f64_16x16x4_f64.
// One workgroup of 8 wavefronts computes the whole NP x NP product of one realization, so Gss and Gs1
// are each read from HBM once (the 64 x 64 tiling reads them twice).  Wavefront w owns row tiles
// {w/2, w/2 + 4} and column tiles {w%2, w%2 + 2, w%2 + 4, w%2 + 6}; at NP = 112 (7 x 7 tiles)
// wavefronts 0..6 own one row tile each with all 7 columns.  k runs in chunks of 16: the next
// chunk is loaded into registers (one complex per 16-byte load) while the current one is multiplied
// out of LDS, then stored to the other LDS buffer -- one barrier per chunk.  b scales the rows of
// Gs1_old in FP64 on the way into LDS.  133 KiB of LDS: one workgroup per CU, 256 registers a lane.
//
// LDS planes are [k][row] (A, transposed on the way in) and [k][col] (B), row length NP + 2 doubles,
// so that the 16 k values of one row written by 16 lanes fall in distinct bank pairs.
// MFMA lane layout (v_mfma_f64_16x16x4_f64): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// result register v of lane l holds D[4v + l/16][l%16].
#include "rg_internal.h"

namespace {

typedef double dy64_d4 __attribute__((ext_vector_type(4)));

template <int NP>
struct DysonShared64 {
    static constexpr int S = NP + 2;
    double a[2][2][16][S];  // [buffer][re, im][k][row]
    double b[2][2][16][S];  // [buffer][re, im][k][col]
};

// G3: Gauss's form of the complex product, P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi),
// Re = P1 - P2, Im = P3 - P1 - P2: three real products on three accumulators instead of four
// products on two (the error of Im is bounded by eps (|P1| + |P2| + |P3|) sums).
// A workgroup computes an MB x MB block of one realization's product: the whole of it (MB = NP) up to
// NP = 128, one of the four NP/2 x NP/2 quadrants at NP = 160, 192, 224, 256 (workgroup 4 r + q: rows
// NP/2 (q / 2) .., columns NP/2 (q % 2) ..), so that each operand element is read twice instead of four times
// by 64 x 64 tiles (NP 160 and 224: 3.38 and 7.24 ms per launch of 4096 against 4.20 and 10.13 ms).
template <int NP, int MB, bool G3>
__global__ __launch_bounds__(512, 2) void k_dyson_gemm_wg64(const double2 *__restrict__ gss,
                                                            const double2 *__restrict__ gs1, double2 *__restrict__ out,
                                                            const double2 *__restrict__ bvec, int n, int cs)
{
    static_assert(NP % MB == 0 && MB % 16 == 0 && MB <= 128, "8 x 8 tiles at most");
    constexpr int QD = NP / MB, QI = QD * QD;                // blocks per dimension and per realization
    constexpr int T = MB / 16, NK = NP / 16, NQ = MB * 16;  // tiles, k chunks, complex per operand chunk
    constexpr int LOADS = (NQ + 511) / 512;
    const int rz = blockIdx.x / QI, row0 = MB * ((blockIdx.x % QI) / QD), col0 = MB * ((blockIdx.x % QI) % QD);
    bvec += (size_t)rz * cs;  // cs = np: constants of each realization, else shared
    __shared__ DysonShared64<MB> sh;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1, li = lane & 15, lg = lane >> 4;
    // ownership of the T x T output tiles: 2 row x 4 column tiles per wavefront, or at T = 7 (49
    // tiles) one row tile and all 7 columns for wavefronts 0..6, which evens the work out (7 tiles at
    // most instead of 8)
    constexpr bool ROW1 = T == 7;
    constexpr int XR = ROW1 ? 1 : 2, YC = ROW1 ? 7 : 4;
    auto rtile = [&](int x) { return ROW1 ? w : wr + 4 * x; };
    auto ctile = [&](int y) { return ROW1 ? y : wc + 2 * y; };
    const size_t base = (size_t)rz * NP * NP;
    const double2 *A = gss + base + (size_t)row0 * NP, *B = gs1 + base + col0;

    double2 ra[LOADS], rb[LOADS];
    // chunk kc: A row idx >> 4, k idx & 15 (16 k of a row contiguous); B k row idx / NP, column idx % NP
    auto load = [&](int kc) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) {
                ra[e] = A[(size_t)(idx >> 4) * NP + kc * 16 + (idx & 15)];
                rb[e] = B[(size_t)(kc * 16 + idx / MB) * NP + idx % MB];
            }
        }
    };
    auto store = [&](int kc, int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) {
                const int row = idx >> 4, k = idx & 15;
                sh.a[buf][0][k][row] = ra[e].x;
                sh.a[buf][1][k][row] = ra[e].y;
                const int kk = idx / MB, col = idx % MB, gk = kc * 16 + kk;
                const double2 bk = gk < n ? bvec[gk] : cmk(0.0, 0.0);
                sh.b[buf][0][kk][col] = bk.x * rb[e].x - bk.y * rb[e].y;
                sh.b[buf][1][kk][col] = bk.x * rb[e].y + bk.y * rb[e].x;
            }
        }
    };
    // G3 stages the next chunk's A and B one after the other (A loaded before the first two k-steps
    // and stored after them, then B), so that one operand's registers are in flight at a time; the
    // other buffer was last read before the previous barrier.
    auto load_a = [&](int kc) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) ra[e] = A[(size_t)(idx >> 4) * NP + kc * 16 + (idx & 15)];
        }
    };
    auto load_b = [&](int kc) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) rb[e] = B[(size_t)(kc * 16 + idx / MB) * NP + idx % MB];
        }
    };
    auto store_a = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) {
                sh.a[buf][0][idx & 15][idx >> 4] = ra[e].x;
                sh.a[buf][1][idx & 15][idx >> 4] = ra[e].y;
            }
        }
    };
    auto store_b = [&](int kc, int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < LOADS; e++) {
            const int idx = tid + 512 * e;
            if (idx < NQ) {
                const int kk = idx / MB, col = idx % MB, gk = kc * 16 + kk;
                const double2 bk = gk < n ? bvec[gk] : cmk(0.0, 0.0);
                sh.b[buf][0][kk][col] = bk.x * rb[e].x - bk.y * rb[e].y;
                sh.b[buf][1][kk][col] = bk.x * rb[e].y + bk.y * rb[e].x;
            }
        }
    };

    // cr, ci: the real and imaginary parts, or P1 and P2 with P3 in cp3 (G3)
    dy64_d4 cr[XR][YC], ci[XR][YC], cp3[G3 ? XR : 1][G3 ? YC : 1];
#pragma unroll
    for (int x = 0; x < XR; x++)
#pragma unroll
        for (int y = 0; y < YC; y++) {
            cr[x][y] = dy64_d4{0.0, 0.0, 0.0, 0.0};
            ci[x][y] = dy64_d4{0.0, 0.0, 0.0, 0.0};
            if constexpr (G3) cp3[x][y] = dy64_d4{0.0, 0.0, 0.0, 0.0};
        }

    load(0);
    store(0, 0);
    __syncthreads();
    for (int kc = 0; kc < NK; kc++) {
        const int buf = kc & 1;
        if (kc + 1 < NK) {
            if constexpr (G3) load_a(kc + 1);
            else load(kc + 1);
        }
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            if constexpr (G3) {
                if (ks == 2 && kc + 1 < NK) {
                    store_a(buf ^ 1);
                    load_b(kc + 1);
                }
                __builtin_amdgcn_sched_barrier(0);  // one k-step's fragments live at a time
            }
            const int kr = 4 * ks + lg;
            double ar[XR], ai[XR], nai[XR], br[YC], bi[YC];
#pragma unroll
            for (int x = 0; x < XR; x++) {
                const int rt = rtile(x);
                ar[x] = rt < T ? sh.a[buf][0][kr][16 * rt + li] : 0.0;
                ai[x] = rt < T ? sh.a[buf][1][kr][16 * rt + li] : 0.0;
                nai[x] = -ai[x];
            }
            if constexpr (G3) {
                // one column tile's operands at a time
#pragma unroll
                for (int y = 0; y < YC; y++) {
                    const int ct = ctile(y);
                    if (ct >= T) continue;  // uniform per wavefront
                    const double b0 = sh.b[buf][0][kr][16 * ct + li], b1 = sh.b[buf][1][kr][16 * ct + li];
#pragma unroll
                    for (int x = 0; x < XR; x++) {
                        if (rtile(x) >= T) continue;
                        cr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], b0, cr[x][y], 0, 0, 0);
                        ci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], b1, ci[x][y], 0, 0, 0);
                        cp3[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x] + ai[x], b0 + b1, cp3[x][y], 0, 0, 0);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                continue;
            }
#pragma unroll
            for (int y = 0; y < YC; y++) {
                const int ct = ctile(y);
                br[y] = ct < T ? sh.b[buf][0][kr][16 * ct + li] : 0.0;
                bi[y] = ct < T ? sh.b[buf][1][kr][16 * ct + li] : 0.0;
            }
#pragma unroll
            for (int x = 0; x < XR; x++)
#pragma unroll
                for (int y = 0; y < YC; y++) {
                    if (rtile(x) >= T || ctile(y) >= T) continue;  // uniform per wavefront
                    cr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], cr[x][y], 0, 0, 0);
                    cr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[x], bi[y], cr[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], ci[x][y], 0, 0, 0);
                    ci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], ci[x][y], 0, 0, 0);
                }
        }
        if (kc + 1 < NK) {
            if constexpr (G3) store_b(kc + 1, buf ^ 1);
            else store(kc + 1, buf ^ 1);
        }
        __syncthreads();
    }
    double2 *C = out + base + (size_t)row0 * NP + col0;
#pragma unroll
    for (int x = 0; x < XR; x++)
#pragma unroll
        for (int y = 0; y < YC; y++) {
            const int rt = rtile(x), ct = ctile(y);
            if (rt >= T || ct >= T) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                double re = cr[x][y][v], im = ci[x][y][v];
                if constexpr (G3) {
                    re = cr[x][y][v] - ci[x][y][v];
                    im = cp3[x][y][v] - cr[x][y][v] - ci[x][y][v];
                }
                C[(size_t)(16 * rt + 4 * v + lg) * NP + 16 * ct + li] = cmk(-re, -im);
            }
        }
}

}  // namespace

bool rg_launch_gemm_wg64(rg_plan *P, int nb, const double2 *gss, const double2 *gs1_old, double2 *gs1_new)
{
    const bool g3 = P->dyson64 == 2;
    switch (P->np) {
#define RG_WG64(N, MB)                                                                                             \
    case N:                                                                                                        \
        if (g3)                                                                                                    \
            hipLaunchKernelGGL((k_dyson_gemm_wg64<N, MB, true>), dim3(nb * (N / MB) * (N / MB)), dim3(512), 0,       \
                               P->stream, gss, gs1_old, gs1_new, P->d_b, P->nmu, P->cstride);                      \
        else                                                                                                       \
            hipLaunchKernelGGL((k_dyson_gemm_wg64<N, MB, false>), dim3(nb * (N / MB) * (N / MB)), dim3(512), 0,      \
                               P->stream, gss, gs1_old, gs1_new, P->d_b, P->nmu, P->cstride);                      \
        return true;
        RG_WG64(80, 80)
        RG_WG64(96, 96)
        RG_WG64(112, 112)
        RG_WG64(128, 128)
        RG_WG64(160, 80)
        RG_WG64(192, 96)
        RG_WG64(224, 112)
        RG_WG64(256, 128)
#undef RG_WG64
        default: return false;
    }
}
