// rg_project_gen.hip -- the disorder projection of large seeded batches with the disorder generated in
// the kernel (gfx950).  The same product as k_project_wide in rg_project.hip, C[b][q] = sum_i U[b][i]
// tab[i][q] over the cells i of one slice, with U the Philox4x32-10 / Box-Muller field of each realization
// (rg_normal_cell); partial sums of the cell ranges to cq[(part nb + b) ntab + q].
//
// Built with MachineLICM off (recurgreen_amd/build.py): hoisted out of the chunk loop, the generator's
// 44 FP64 constants take 88 registers that the accumulators need at three waves per SIMD (spills); left in
// the loop they are scalar moves.
#include "rg_internal.h"

namespace {

constexpr int PR_KC = 16;  // transverse cells per staged chunk
typedef double dbl4 __attribute__((ext_vector_type(4)));

// The projection of large seeded batches with the disorder generated in the kernel: 12 waves per
// workgroup of 64 realizations, three per SIMD.  Wave w multiplies rows 16 (w mod 4).. against column third
// w / 4 of the table (tiles NT g / 3 .. NT (g + 1) / 3 - 1); waves generate the next chunk's 16 cells of the
// 64 realizations into LDS (waves 0..7: the two cells 2 w, 2 w + 1 of one Philox block and Box-Muller pair;
// two pairs per SIMD), which the three thirds read, so each cell is generated once.  The generator's FP64 arithmetic does not overlap the
// FP64 products (they share the FP64 datapath: the vector and matrix FP64 peaks are equal): at 4096 x N=128
// (ny = 1285) the products alone take 62 us per slice, the generator 37-41 us more, in this kernel or with
// the generator inside the 128-realization one alike; this kernel's gain (97 against 101 us) comes from
// 64-realization workgroups, which fill the CUs with 4 cell ranges instead of 8 (half the partial sums
// written and re-read, 3.7% chunk imbalance instead of 8.6%).  Four waves per SIMD leave 128 registers
// per wave: too few for the accumulators beside the generator.
constexpr int PG_RB = 64;
constexpr int PG_SA = PG_RB + 4;  // row stride of the staged disorder: in each half-wave, 64-bit reads of
                                  // rows 4 apart land 32 banks apart

template <int NTG, int NT>
__device__ __forceinline__ void pg_products(dbl4 *acc, const double (*s_a)[PG_SA], const double (*s_t)[16 * NT],
                                            int row, int tile0, int li, int lk)
{
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const double a = s_a[4 * lk + s][row];
#pragma unroll
        for (int t = 0; t < NTG; t++)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, s_t[4 * lk + s][16 * (tile0 + t) + li], acc[t], 0, 0, 0);
    }
}

template <int NT>
__global__ __launch_bounds__(768, 1) void k_project_gen(int ny, int ixd, int64_t seed0,
                                                        const int64_t *__restrict__ seedp, double stdu,
                                                        const int64_t *__restrict__ seeds,
                                                        const double *__restrict__ stdus,
                                                        const double *__restrict__ tab, int ntp, int ntab,
                                                        int nchunk, int parts, double *__restrict__ cq, int nb)
{
    constexpr int NG = (NT + 2) / 3;  // tiles of the largest third (the others: NG or NG - 1)
    __shared__ double s_t[2][PR_KC][16 * NT];
    __shared__ double s_a[2][PR_KC][PG_SA];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
    const int g = __builtin_amdgcn_readfirstlane(w >> 2), row = 16 * (w & 3) + li;
    const int tile0 = NT * g / 3, ntg = NT * (g + 1) / 3 - tile0;
    const int b0 = blockIdx.x * PG_RB, part = blockIdx.y;
    const int c0 = (int)((int64_t)part * nchunk / parts), c1 = (int)((int64_t)(part + 1) * nchunk / parts);
    // the generator's realization (its seed and strength read once; past the batch: the last one's, zeroed)
    const int gb = b0 + lane;
    const bool gok = gb < nb;
    const int gbc = gok ? gb : nb - 1;
    const uint64_t seed = (uint64_t)(seeds ? seeds[gbc] : (seedp ? *seedp + seed0 : seed0) + gbc);
    const double sd = stdus ? stdus[gbc] : stdu;
    // wave w < 8: the cells 2 w, 2 w + 1 of chunk c (one Philox block and Box-Muller pair)
    auto generate = [&](int c, int buf) {
        if (w >= PR_KC / 2) return;  // uniform per wavefront
        const int iy = c * PR_KC + 2 * w;
        // the key schedule derived from the seed is recomputed per chunk rather than hoisted out of the chunk
        // loop, where its registers would stay live beside the accumulators
        uint64_t sj = seed;
        asm volatile("" : "+v"(sj));
        double z0, z1;
        rg_normal_pair(sj, 0u, ixd, iy >> 1, z0, z1);
        // products with the 0/1 masks rather than selects, which the compiler turns into branches
        s_a[buf][2 * w][lane] = (sd * z0) * ((gok && iy < ny) ? 1.0 : 0.0);
        s_a[buf][2 * w + 1][lane] = (sd * z1) * ((gok && iy + 1 < ny) ? 1.0 : 0.0);
    };
    // table rows 16 c .. 16 c + 15 straight into LDS (the 16 bytes of lane l at the wave's LDS base
    // + 16 l, so the staged rows are contiguous), over the 12 waves
    auto stage_tab = [&](int c, int buf) {
        constexpr int RU = 8 * NT, UNITS = PR_KC * RU;
        double *dst = &s_t[buf][0][0];
        for (int r = 0; r < UNITS; r += 768) {
            const int u0 = r + 64 * w;
            if (u0 < UNITS) {
                const int u = u0 + lane, col = 2 * (u % RU);
                const double *src = tab + (size_t)(c * PR_KC + u / RU) * ntp + (col < ntp ? col : 0);
                __builtin_amdgcn_global_load_lds(src, dst + 2 * u0, 16, 0, 0);
            }
        }
    };

    dbl4 acc[NG];
#pragma unroll
    for (int t = 0; t < NG; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    if (c0 < c1) {
        stage_tab(c0, 0);
        generate(c0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int c = c0; c < c1; c++) {
        const int buf = (c - c0) & 1;
        if (c + 1 < c1) {
            stage_tab(c + 1, buf ^ 1);
            generate(c + 1, buf ^ 1);
        }
        if (ntg == NG)
            pg_products<NG, NT>(acc, s_a[buf], s_t[buf], row, tile0, li, lk);
        else
            pg_products<NG - 1, NT>(acc, s_a[buf], s_t[buf], row, tile0, li, lk);
        __builtin_amdgcn_s_waitcnt(0);  // the next chunk's table rows are in LDS
        __syncthreads();
    }
    double *out = cq + (size_t)part * nb * ntab;
#pragma unroll
    for (int t = 0; t < NG; t++) {
        if (t >= ntg) break;
        const int col = 16 * (tile0 + t) + li;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int bb = b0 + 16 * (w & 3) + 4 * v + lk;
            if (bb < nb && col < ntab) out[(size_t)bb * ntab + col] = acc[t][v];
        }
    }
}

}  // namespace

bool rg_launch_project_gen(rg_plan *P, int nb, int ixd, int64_t seed0)
{
    const int nti = P->ntp / 16;
    if (nti < 3 || nti > 17) return false;
    // Cell ranges (partial sums): one workgroup per CU, so the fewest that minimize rounds of workgroups x
    // chunks per workgroup (4096 realizations: 64 workgroups x 4 parts, 21 chunks each at ny = 1285, against
    // 8 parts of 128-realization workgroups in k_project_wide: half the partial sums to write and re-read).
    const int rb = (nb + PG_RB - 1) / PG_RB;
    int parts = 1;
    int64_t best = -1;
    for (int p = 1; p <= P->cq_parts; p++) {
        const int64_t cost = (int64_t)((rb * p + P->ncu - 1) / P->ncu) * ((P->nchunk + p - 1) / p);
        if (best < 0 || cost < best) {
            best = cost;
            parts = p;
        }
    }
    P->cq_parts_used = parts;
    const dim3 grid(rb, parts);
    switch (nti) {
#define RG_PG(T)                                                                                                   \
    case T:                                                                                                        \
        hipLaunchKernelGGL((k_project_gen<T>), grid, dim3(768), 0, P->stream, P->ny, ixd, seed0, P->seed_dev,       \
                           P->stdu, P->d_seeds_r, P->d_stdu_r, P->d_tab, P->ntp, P->ntab, P->nchunk, parts, P->d_cq, \
                           nb);                                                                                    \
        return true;
        RG_PG(3) RG_PG(4) RG_PG(5) RG_PG(6) RG_PG(7) RG_PG(8) RG_PG(9)
        RG_PG(10) RG_PG(11) RG_PG(12) RG_PG(13) RG_PG(14) RG_PG(15) RG_PG(16) RG_PG(17)
#undef RG_PG
        default: return false;
    }
}
