// rg_project.hip -- disorder slice projection on the FP64 matrix cores (gfx950).
//
// src/waveguide.f90:366-380 adds modal^H diag(u_x) modal to every disordered slice.  With sine
// modes (Dirichlet) chi_m chi_n = (1/ny)[cos((m-n) pi y) - cos((m+n) pi y)], so the projected slice
// is V_mn = C(|m-n|) - C(m+n+2) (0-based m, n) with C(q) = (1/ny) sum_i u_i cos(q pi y_i); with
// plane waves (periodic) V_mn = D(n-m), D(q) = C(2q) + i C(2q+1) over a cos/sin table.  For the
// whole batch and one slice this is the product  C[b][q] = sum_i U[b][i] tab[i][q]:
// M = realizations, N = table columns (2N+1 or 2N), K = ny transverse cells.
//
// Batches of at least the plan's project_wide_min realizations (default 1: all) use k_project_gen
// (rg_project_gen.hip: seeded, up to 128 modes, the disorder generated in the kernel) or k_project_wide
// (below: 128 realizations per workgroup, the table staged straight into LDS).  Smaller ones take
// k_project_mfma: each workgroup (4 wavefronts) owns 32 realizations (two 16-row MFMA tiles), one
// contiguous range of 16-cell chunks of K and one group of up to 20 table column tiles (blockIdx.z);
// wavefront w owns the tiles {w, w+4, ...} of its group.  The slice
// values U (generated for the slice by k_noise_slice, or read from an explicit potential batch) are
// staged through LDS one chunk ahead of the v_mfma_f64_16x16x4 products; table tiles stream from
// L2.
// Partial sums of the K ranges go to cq[(part * nb + b) * ntab + q]; the slice-inversion kernels add
// them in part order and scale by 1/ny, so the result does not depend on scheduling.
//
// MFMA lane layout (v_mfma_f64_16x16x4_f64, measured): A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, result register v of lane l holds D[4v + l/16][l%16].
#include "rg_internal.h"

namespace {

constexpr int PR_RB = 32;     // realizations per workgroup
constexpr int PR_KC = 16;     // transverse cells per staged chunk
constexpr int PR_MAXNT = 20;  // table column tiles supported (ntab <= 320)
constexpr int PR_TPW = (PR_MAXNT + 3) / 4;

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Disorder of realization b, cells 2 p and 2 p + 1 of slice ixd: stdu N(seed, ixd, 2 p + {0, 1}) with the seed
// seed0 + b (or *seedp + seed0 + b in a captured sweep, or seeds[b] and stdus[b] in an uneven batch).
__device__ __forceinline__ void noise_pair(int b, int ixd, int p, int64_t seed0, const int64_t *__restrict__ seedp,
                                           double stdu, const int64_t *__restrict__ seeds,
                                           const double *__restrict__ stdus, double &z0, double &z1)
{
    const int64_t s = seeds ? seeds[b] : (seedp ? *seedp + seed0 : seed0) + b;
    const double sd = stdus ? stdus[b] : stdu;
    rg_normal_pair((uint64_t)s, 0u, ixd, p, z0, z1);
    z0 *= sd;
    z1 *= sd;
}

// One chunk of slice values into s_u[KC][RB]: two cells per thread, u(b, iy) = us[b sb + iy si].
__device__ inline void stage_chunk(double (*s_u)[PR_RB + 1], int chunk, int b0, int nb, const double *__restrict__ us,
                                   int64_t sb, int si, int ny, int tid)
{
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int e = tid + 256 * h;
        const int rr = e & (PR_RB - 1), ii = e >> 5;
        const int b = b0 + rr, iy = chunk * PR_KC + ii;
        s_u[ii][rr] = (b < nb && iy < ny) ? us[(size_t)b * sb + (size_t)iy * si] : 0.0;
    }
}

// Disorder of one slice for a batch of seeds: us[b ny + iy] = stdu * N(seed0 + b, ixd, iy), the
// Philox4x32-10 / Box-Muller normals of the cells, two per thread (one block and Box-Muller pair; generated
// here, at full occupancy, so that the projection kernel's matrix cores are not held up by the
// transcendental functions).  With seedp != null the first seed is *seedp + seed0 (a captured sweep reads it
// from device memory).  An uneven batch gives every realization its own seed and strength: seeds[b],
// stdus[b].  Grid: (ceil(ny / 512), nb).
__global__ __launch_bounds__(256) void k_noise_slice(int ixd, int64_t seed0, const int64_t *__restrict__ seedp, int ny,
                                                     double stdu, const int64_t *__restrict__ seeds,
                                                     const double *__restrict__ stdus, double *__restrict__ us)
{
    const int b = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x, iy = 2 * p;
    if (iy >= ny) return;
    double z0, z1;
    noise_pair(b, ixd, p, seed0, seedp, stdu, seeds, stdus, z0, z1);
    us[(size_t)b * ny + iy] = z0;
    if (iy + 1 < ny) us[(size_t)b * ny + iy + 1] = z1;
}

// The projection of large batches from slice values in memory (explicit potentials, or the disorder of
// k_noise_slice): each workgroup (8 wavefronts) owns 128 realizations, one contiguous range of 16-cell
// chunks and one group of NT column tiles (blockIdx.z: tiles NT z ..; above 128 modes the table's 33 tiles
// form two groups); wavefront w computes rows 16w.. against the group's tiles.  The table chunk
// (16 x 16 NT doubles) is staged once per workgroup in LDS, one chunk ahead, so the table is read from L2
// once per 128 realizations; the slice values of the next chunk are loaded while the current chunk's
// products run.  MFMA k-step s of a chunk takes cell 4 (lane / 16) + s of it in lane group lane / 16, for A
// and B alike.  Tiles past the table's last (the second group's 17th) are staged from a valid address and
// their sums are not written.  (Seeded batches of one column group: k_project_gen, rg_project_gen.hip.)
constexpr int PW_RB = 128;
constexpr int PW_MAXNT = 17;  // column tiles per group (ntab <= 272 in one group: up to 128 modes)

template <int NT>
__global__ __launch_bounds__(512, 1) void k_project_wide(const double *__restrict__ us, int64_t sb, int si, int ny,
                                                         const double *__restrict__ tab, int ntp, int ntab,
                                                         int nchunk, int parts, double *__restrict__ cq, int nb)
{
    __shared__ double s_t[2][PR_KC][16 * NT];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
    const int b0 = blockIdx.x * PW_RB, part = blockIdx.y, tile0 = NT * blockIdx.z;
    const int c0 = (int)((int64_t)part * nchunk / parts), c1 = (int)((int64_t)(part + 1) * nchunk / parts);
    const int b = b0 + 16 * w + li;  // the realization of this lane's A operand rows
    const bool bok = b < nb;

    dbl4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    double a_cur[4], a_nxt[4];
    auto a_chunk = [&](int c, double (&a)[4]) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int iy = c * PR_KC + 4 * lk + s;
            a[s] = 0.0;
            if (bok && iy < ny) a[s] = us[(size_t)b * sb + (size_t)iy * si];
        }
    };
    // Rows 16 c .. 16 c + 15, columns 16 tile0 .. 16 (tile0 + NT) - 1 of the zero-padded table (ntp columns),
    // loaded straight into LDS: lane l of a wave writes the 16 bytes at the wave's LDS base + 16 l, so the
    // staged rows are contiguous in LDS whatever their stride in the table.
    auto stage_tab = [&](int c, int buf) {
        constexpr int RU = 8 * NT, UNITS = PR_KC * RU;  // 16-byte units per staged row and per chunk (x 64)
        double *dst = &s_t[buf][0][0];
        for (int r = 0; r < UNITS; r += 512) {
            const int u0 = r + 64 * w;  // this wave's first unit (wave-uniform test)
            if (u0 < UNITS) {
                const int u = u0 + lane, col = 16 * tile0 + 2 * (u % RU);
                const double *src = tab + (size_t)(c * PR_KC + u / RU) * ntp + (col < ntp ? col : 0);
                __builtin_amdgcn_global_load_lds(src, dst + 2 * u0, 16, 0, 0);
            }
        }
    };
    if (c0 < c1) {
        stage_tab(c0, 0);
        a_chunk(c0, a_cur);
    }
    __builtin_amdgcn_s_waitcnt(0);  // the staged chunk is in LDS
    __syncthreads();
    for (int c = c0; c < c1; c++) {
        const int buf = (c - c0) & 1;
        if (c + 1 < c1) {
            stage_tab(c + 1, buf ^ 1);
            a_chunk(c + 1, a_nxt);
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_cur[s], s_t[buf][4 * lk + s][16 * t + li], acc[t], 0, 0, 0);
        __builtin_amdgcn_s_waitcnt(0);  // the next chunk is in LDS
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; s++) a_cur[s] = a_nxt[s];
    }
    double *out = cq + (size_t)part * nb * ntab;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int q = 16 * (tile0 + t) + li;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int bb = b0 + 16 * w + 4 * v + lk;
            if (bb < nb && q < ntab) out[(size_t)bb * ntab + q] = acc[t][v];
        }
    }
}

__global__ __launch_bounds__(256) void k_project_mfma(const double *__restrict__ us, int64_t sb, int si, int ny,
                                                      const double *__restrict__ tab, int ntp, int ntab,
                                                         int nchunk, int parts, double *__restrict__ cq, int nb)
{
    __shared__ double s_u[2][PR_KC][PR_RB + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b0 = blockIdx.x * PR_RB, part = blockIdx.y, nt0 = PR_MAXNT * blockIdx.z;  // column tile group
    const int c0 = (int)((int64_t)part * nchunk / parts), c1 = (int)((int64_t)(part + 1) * nchunk / parts);
    const int ntiles = ntp / 16 < nt0 + PR_MAXNT ? ntp / 16 : nt0 + PR_MAXNT;
    const int li = lane & 15, lk = lane >> 4;

    dbl4 acc[2][PR_TPW];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int j = 0; j < PR_TPW; j++) acc[m][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    if (c0 < c1) stage_chunk(s_u[0], c0, b0, nb, us, sb, si, ny, tid);
    __syncthreads();
    for (int c = c0; c < c1; c++) {
        const int buf = (c - c0) & 1;
        if (c + 1 < c1) stage_chunk(s_u[buf ^ 1], c + 1, b0, nb, us, sb, si, ny, tid);
        const double *trow = tab + (size_t)(c * PR_KC + lk) * ntp + li;
#pragma unroll
        for (int kk = 0; kk < PR_KC / 4; kk++) {
            const double a0 = s_u[buf][4 * kk + lk][li];
            const double a1 = s_u[buf][4 * kk + lk][16 + li];
#pragma unroll
            for (int j = 0; j < PR_TPW; j++) {
                const int nt = nt0 + w + 4 * j;
                if (nt < ntiles) {
                    const double bv = trow[(size_t)(4 * kk) * ntp + 16 * nt];
                    acc[0][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv, acc[0][j], 0, 0, 0);
                    acc[1][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv, acc[1][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    double *out = cq + (size_t)part * nb * ntab;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int j = 0; j < PR_TPW; j++) {
            const int nt = nt0 + w + 4 * j;
            if (nt >= ntiles) continue;
            const int q = 16 * nt + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int b = b0 + 16 * m + 4 * v + lk;
                if (b < nb && q < ntab) out[(size_t)b * ntab + q] = acc[m][j][v];
            }
        }
}

}  // namespace

int rg_project_parts(int ny)
{
    const int nchunk = (ny + PR_KC - 1) / PR_KC;
    int parts = (nchunk + 9) / 10;  // about ten chunks per workgroup
    return parts < 1 ? 1 : (parts > 8 ? 8 : parts);
}

bool rg_project_supported(int ntab) { return (ntab + 15) / 16 <= 2 * PR_MAXNT; }

void rg_launch_project(rg_plan *P, int nb, int ixd, int64_t seed0, const double *d_upot, int upot_nreal)
{
    (void)upot_nreal;
    P->cq_parts_used = P->cq_parts;
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_PROJECT, &stop);
    const double *us = d_upot ? d_upot + ixd : nullptr;  // explicit potentials upot(ix, iy, b)
    int64_t sb = (int64_t)P->nxdiso * P->ny;
    int si = P->nxdiso;
    // Large batches (one workgroup per CU and partial sum at least): seeded batches of one column group
    // generate their disorder in k_project_gen unless RG_PROJECT_GEN=0; the others run k_project_wide.
    static const int gen_env = [] {
        const char *e = std::getenv("RG_PROJECT_GEN");
        return e ? std::atoi(e) : 1;
    }();
    const int nti = P->ntp / 16;
    // one group of column tiles (up to 128 modes), or two (up to 256 modes, the disorder from k_noise_slice
    // rather than generated twice: 0.52 -> 0.30 ms per slice of 4096 realizations at N = 256, ny = 2565),
    // for every batch of at least project_wide_min realizations (default 1): about twice as fast as
    // k_noise_slice + k_project_mfma from 128 realizations up (N = 128, 256 realizations: 50 against 90 us
    // per slice; N = 256, 512: 119 against 224 us)
    const int tgroups = (nti + PW_MAXNT - 1) / PW_MAXNT;
    if (tgroups <= 2 && nb >= P->project_wide_min) {
        if (!d_upot && gen_env && tgroups == 1 && rg_launch_project_gen(P, nb, ixd, seed0)) {
            rg_timer_end(P, RGK_PROJECT, stop);
            return;
        }
        if (!d_upot) {
            hipLaunchKernelGGL(k_noise_slice, dim3((P->ny + 511) / 512, nb), dim3(256), 0, P->stream, ixd, seed0,
                               P->seed_dev, P->ny, P->stdu, P->d_seeds_r, P->d_stdu_r, P->d_us);
            us = P->d_us;
            sb = P->ny;
            si = 1;
        }
        const dim3 grid((nb + PW_RB - 1) / PW_RB, P->cq_parts, tgroups);
        switch (tgroups == 1 ? nti : PW_MAXNT) {
#define RG_PW(T)                                                                                                   \
    case T:                                                                                                        \
        hipLaunchKernelGGL(k_project_wide<T>, grid, dim3(512), 0, P->stream, us, sb, si, P->ny, P->d_tab, P->ntp,   \
                           P->ntab, P->nchunk, P->cq_parts, P->d_cq, nb);                                          \
        break;
            RG_PW(1) RG_PW(2) RG_PW(3) RG_PW(4) RG_PW(5) RG_PW(6) RG_PW(7) RG_PW(8) RG_PW(9)
            RG_PW(10) RG_PW(11) RG_PW(12) RG_PW(13) RG_PW(14) RG_PW(15) RG_PW(16) RG_PW(17)
#undef RG_PW
            default: break;
        }
        rg_timer_end(P, RGK_PROJECT, stop);
        return;
    }
    if (!d_upot) {  // seeded: this slice's disorder first
        hipLaunchKernelGGL(k_noise_slice, dim3((P->ny + 511) / 512, nb), dim3(256), 0, P->stream, ixd, seed0,
                           P->seed_dev, P->ny, P->stdu, P->d_seeds_r, P->d_stdu_r, P->d_us);
        us = P->d_us;
        sb = P->ny;
        si = 1;
    }
    const int groups = (P->ntp / 16 + PR_MAXNT - 1) / PR_MAXNT;  // column tiles in groups of PR_MAXNT
    hipLaunchKernelGGL(k_project_mfma, dim3((nb + PR_RB - 1) / PR_RB, P->cq_parts, groups), dim3(256), 0, P->stream, us, sb, si,
                       P->ny, P->d_tab, P->ntp, P->ntab, P->nchunk, P->cq_parts, P->d_cq, nb);
    rg_timer_end(P, RGK_PROJECT, stop);
}

rg_cqv rg_cq_view(const rg_plan *P, int nb, bool with_disorder)
{
    rg_cqv v;
    v.cq = with_disorder ? P->d_cq : nullptr;
    v.ntab = P->ntab;
    v.nb = nb;
    v.parts = P->cq_parts_used;
    v.scale = 1.0 / P->ny;
    return v;
}
