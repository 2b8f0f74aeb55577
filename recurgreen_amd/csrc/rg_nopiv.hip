// rg_nopiv.hip -- blocked Gauss-Jordan slice inversion with verified diagonal pivots, FP32 state on
// the matrix cores (gfx950).
//
// Slice inversion of the mixed sweeps, src/waveguide.f90:434-435 (zgetrf + zgetri of
// src/base_utils.f90:235-274).  zgetrf pivots on max |Re|+|Im| of the current column over the
// rows not pivoted yet.  On the slice matrices of the sweep this choice is the diagonal element
// (the RGF Schur complements A_x - B Gss B), so the elimination here takes the diagonal pivots in
// 16-column blocks on the matrix cores and VERIFIES, with FP64 comparisons, that every pivot is the
// one zgetrf's rule picks (within a relative margin of 1e-4): inside a block from the elimination
// itself, below it from V = P U_D^-1 diag(u) (P: the rows below the block, D = L_D U_D the block,
// u its pivots; V_ij is the candidate entry of row i at step j).  A realization whose check fails
// (or meets a zero pivot) is not written: it is flagged in `redo` and inverted again by the
// pivoted kernel (k_gj_lookahead), which starts from the same, unmodified Gss.
//
// Layout: wavefront w owns the 16-column block w as np/16 MFMA accumulator tiles of 16x16 (real
// and imaginary planes), register v of lane l holding row 4(l/16) + v, column l%16
// (v_mfma_f32_16x16x4_f32).  Block step K (D = tile (K,K), P = column block K):
//   wavefront K: Gauss-Jordan on D in registers (16 unrolled steps, cross-lane broadcasts by
//   ds_bpermute and DPP, pivot checks after the last step) -> D^-1, pivots, and
//   W = D^-1 L_D diag(u) = U_D^-1 diag(u); writes P (rows outside K), D^-1 and W to LDS;
//   every wavefront: Rt = D^-1 X[K, w] (its own tile K), X[K, w] := Rt, X[i, w] -= P_i Rt (i != K)
//   (wavefront K: X[i, K] := -P_i D^-1, X[K, K] := D^-1), and checks the rows of its own tile
//   index against V.  The product operands use the accumulator tiles directly: a k-step v of the
//   MFMA takes k = 4g + v in lane group g, so a tile in accumulator layout is a valid B operand.
// Look-ahead: wavefront K+1 updates its diagonal tile first and inverts the next D before its
// other tiles, so one workgroup barrier per block step.
#include "rg_internal.h"
#include <utility>

namespace {

typedef float np_f4 __attribute__((ext_vector_type(4)));
constexpr double NP_MARGIN = 1.0 - 1e-4;
constexpr int LDW = 20;  // LDS row stride (floats): 16-byte aligned rows, four k values per read

template <int WV>
struct NopivShared {
    static constexpr int NP = 16 * WV;
    alignas(16) float pr[2][NP][LDW], pi[2][NP][LDW];  // P (column block, rows outside the block; zero inside)
    alignas(16) float dr[2][16][LDW], di[2][16][LDW];  // D^-1, row-major
    alignas(16) float dtr[2][16][LDW], dti[2][16][LDW];  // D^-1, transposed (B operand reads)
    alignas(16) float wtr[2][16][LDW], wti[2][16][LDW];  // W = U_D^-1 diag(u), transposed
    double ua[2][16];                    // |u_j|_1 (FP64)
    int bad;
    alignas(16) double cq[2 * 256 + 8];
    double2 bv[NP], ad[NP];  // b and the slice diagonal
};

__device__ inline float readlane_f(float v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); }

// Lane (c, g) <- lane (J, g): DPP row_newbcast:J.
template <int J>
__device__ __attribute__((always_inline)) inline float lane_bcast(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x150 + J, 0xF, 0xF, true));
}

// Step J of the Gauss-Jordan elimination of a 16x16 block (xr, xi) held in accumulator layout by
// one wavefront, diagonal pivot.  The pivot comes from its lane (readlane), the pivot row from a
// row broadcast, the pivot column from a DPP lane broadcast: all register-to-register, with J a
// constant.  Lane c == J records the pivot magnitude and the reduced column (rows >= J) for the
// checks and for L_D diag(u).
template <int J>
__device__ __attribute__((always_inline)) inline void gj_step(np_f4 &xr, np_f4 &xi, np_f4 &lr, np_f4 &li, double &my_ua,
                                                              bool &my_den, int lane)
{
    constexpr int VJ = J & 3, G = J >> 2;
    const int c = lane & 15, g = lane >> 4;
    // pivot D[J][J] from its lane; pivot row D[J][c]; column J of my rows D[4g+v][J]
    const float pvr = readlane_f(xr[VJ], J + 16 * G), pvi = readlane_f(xi[VJ], J + 16 * G);
    const float prr = __shfl(xr[VJ], c + 16 * G, 64), pri = __shfl(xi[VJ], c + 16 * G, 64);
    float mr[4], mi[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        mr[v] = lane_bcast<J>(xr[v]);
        mi[v] = lane_bcast<J>(xi[v]);
    }
    const float den = pvr * pvr + pvi * pvi;
    // 1/den: v_rcp_f32 (1 ulp) and one Newton step, which removes its one-sided error (a
    // per-slice bias that would otherwise build up over the sweep) at 2 FMAs of latency
    float inv = 1.0f;
    if (den > 0.f) {
        const float r0 = __builtin_amdgcn_rcpf(den);
        inv = fmaf(r0, fmaf(-den, r0, 1.0f), r0);
    }
    const float dr = pvr * inv, di = -pvi * inv;
    if (c == J) {
        my_ua = fabs((double)pvr) + fabs((double)pvi);
        my_den = den > 0.f;
    }
    if (c == J) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 4 * g + v;
            lr[v] = row >= J ? mr[v] : 0.f;
            li[v] = row >= J ? mi[v] : 0.f;
        }
    }
    asm volatile("" : "+v"(lr), "+v"(li));
    // scaled pivot row (column J of X reset to e_J first, so its entry is d)
    float rvr, rvi;
    if (c == J) { rvr = dr; rvi = di; }
    else { rvr = prr * dr - pri * di; rvi = prr * di + pri * dr; }
    const float keep = c == J ? 0.0f : 1.0f;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        float mur = mr[v], mui = mi[v];
        if (v == VJ && g == G) { mur = pvr - 1.0f; mui = pvi; }
        float nr = fmaf(mui, rvi, -(mur * rvr));
        const float ni = fmaf(-mur, rvi, -(mui * rvr));
        if (v == VJ) nr += (g == G && c == J) ? 1.0f : 0.0f;
        xr[v] = fmaf(keep, xr[v], nr);
        xi[v] = fmaf(keep, xi[v], ni);
    }
}

template <int... Js>
__device__ __attribute__((always_inline)) inline void gj_steps(std::integer_sequence<int, Js...>, np_f4 &xr, np_f4 &xi,
                                                               np_f4 &lr, np_f4 &li, double &my_ua, bool &my_den,
                                                               int lane)
{
    (gj_step<Js>(xr, xi, lr, li, my_ua, my_den, lane), ...);
}

// Gauss-Jordan on the 16x16 block, 16 steps unrolled.  On return (xr, xi) = D^-1, (lr, li) =
// L_D diag(u) (the column of the reduced block at each step, rows >= step), ua[j] = |u_j|_1;
// returns false if a pivot fails the zgetrf check (FP64 comparisons) or is zero.
__device__ __attribute__((always_inline)) inline bool block_gj(np_f4 &xr, np_f4 &xi, np_f4 &lr, np_f4 &li, double *ua,
                                                               int lane)
{
    const int c = lane & 15, g = lane >> 4;
    lr = np_f4{0.f, 0.f, 0.f, 0.f};
    li = np_f4{0.f, 0.f, 0.f, 0.f};
    double my_ua = 0.0;  // lane c keeps |u_c|_1
    bool my_den = true;  // and whether |u_c|^2 > 0 in FP32
    // the per-step lane masks are recomputed here rather than hoisted out of the caller's loop
    int ln = lane;
    asm volatile("" : "+v"(ln));
    gj_steps(std::make_integer_sequence<int, 16>{}, xr, xi, lr, li, my_ua, my_den, ln);
    // The pivot checks of all 16 steps at once: lane c holds the candidates of step c (column c
    // of the reduced block, rows >= c, in lr/li) and |u_c|_1; a row below the pivot whose entry
    // exceeds the margin means zgetrf would have picked another pivot.
    bool ok = my_ua > 0.0 && my_den;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int row = 4 * g + v;
        const bool big = fabs((double)lr[v]) + fabs((double)li[v]) > NP_MARGIN * my_ua;
        ok = ok && !(row > c && big);
    }
    if (lane < 16) ua[c] = my_ua;
    // every lane must agree: combine the flags over the wavefront
    return __all(ok);
}

// C (tile) -= A B for one 16x16 complex tile: A rows from LDS ([row][k], k = block column, four
// consecutive k per lane in one read), B in accumulator layout (k-step v = rows 4g + v).
template <int S>
__device__ __attribute__((always_inline)) inline void tile_fnms(np_f4 &cr, np_f4 &ci, const float (*asr)[S],
                                                                const float (*asi)[S], int r0, np_f4 br, np_f4 bi,
                                                                int lane)
{
    const int i = lane & 15, g = lane >> 4;
    const np_f4 a_r = *reinterpret_cast<const np_f4 *>(&asr[r0 + i][4 * g]);
    const np_f4 a_i = *reinterpret_cast<const np_f4 *>(&asi[r0 + i][4 * g]);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_r[v], br[v], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a_i[v], bi[v], cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_r[v], bi[v], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_i[v], br[v], ci, 0, 0, 0);
    }
}

// C = A B for A from LDS [16][LDW] (rows i, k = 4g+v) and B in accumulator layout.
template <int S>
__device__ __attribute__((always_inline)) inline void tile_mul(np_f4 &cr, np_f4 &ci, const float (*asr)[S],
                                                               const float (*asi)[S], np_f4 br, np_f4 bi, int lane)
{
    const int i = lane & 15, g = lane >> 4;
    const np_f4 a_r = *reinterpret_cast<const np_f4 *>(&asr[i][4 * g]);
    const np_f4 a_i = *reinterpret_cast<const np_f4 *>(&asi[i][4 * g]);
    cr = np_f4{0.f, 0.f, 0.f, 0.f};
    ci = np_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r[v], br[v], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_i[v], bi[v], cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r[v], bi[v], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(a_i[v], br[v], ci, 0, 0, 0);
    }
}

// B operand (accumulator layout) from a transposed LDS 16x16 matrix St[col][row]: reg v of lane l
// = B[4g+v][l%16] = St[l%16][4g+v], one read.
template <int S>
__device__ inline void load_b(np_f4 &br, np_f4 &bi, const float (*tr)[S], const float (*ti)[S], int lane)
{
    const int c = lane & 15, g = lane >> 4;
    br = *reinterpret_cast<const np_f4 *>(&tr[c][4 * g]);
    bi = *reinterpret_cast<const np_f4 *>(&ti[c][4 * g]);
}

// The owner of block K (wavefront K) inverts its diagonal tile, held in accumulator slot SLOT, and
// publishes D^-1, W and the pivot magnitudes to buffer K & 1.
template <int WV, int SLOT>
__device__ __attribute__((always_inline)) inline void publish_inverse(const np_f4 (&ar)[WV], const np_f4 (&ai)[WV],
                                                                      int K, int lane, NopivShared<WV> &sh)
{
    const int c = lane & 15, g = lane >> 4, buf = K & 1;
    np_f4 dr = ar[SLOT], di = ai[SLOT], lr, li;
    const bool ok = block_gj(dr, di, lr, li, sh.ua[buf], lane);
    if (!ok && lane == 0) sh.bad = 1;
#pragma unroll
    for (int v = 0; v < 4; v++) {  // D^-1 to LDS, row-major and transposed
        sh.dr[buf][4 * g + v][c] = dr[v];
        sh.di[buf][4 * g + v][c] = di[v];
    }
    *reinterpret_cast<np_f4 *>(&sh.dtr[buf][c][4 * g]) = dr;
    *reinterpret_cast<np_f4 *>(&sh.dti[buf][c][4 * g]) = di;
    __builtin_amdgcn_wave_barrier();
    np_f4 wr, wi;
    tile_mul(wr, wi, sh.dr[buf], sh.di[buf], lr, li, lane);  // W = D^-1 (L_D diag(u))
    *reinterpret_cast<np_f4 *>(&sh.wtr[buf][c][4 * g]) = wr;
    *reinterpret_cast<np_f4 *>(&sh.wti[buf][c][4 * g]) = wi;
}

// ... and the column block P of block K (rows outside the block; zero inside) to buffer K & 1,
// once every slot holds its updated tile.  Slot s holds row block (s + rot) mod WV.
template <int WV, int SLOT>
__device__ __attribute__((always_inline)) inline void publish_column(const np_f4 (&ar)[WV], const np_f4 (&ai)[WV],
                                                                     int K, int rot, int lane, NopivShared<WV> &sh)
{
    const int c = lane & 15, g = lane >> 4, buf = K & 1;
#pragma unroll
    for (int s = 0; s < WV; s++) {
        const int rb = (s + rot) % WV;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            sh.pr[buf][16 * rb + 4 * g + v][c] = s == SLOT ? 0.f : ar[s][v];
            sh.pi[buf][16 * rb + 4 * g + v][c] = s == SLOT ? 0.f : ai[s][v];
        }
    }
}

template <int WV>
__global__ __launch_bounds__(64 * WV, 4) void k_inverse_nopiv(float2 *__restrict__ gss, const rg_cqv cqv, int boundtype,
                                                              const double2 *__restrict__ adiag,
                                                              const double2 *__restrict__ bvec, int n, int first,
                                                              int32_t *__restrict__ redo, int cs,
                                                              const int32_t *__restrict__ nxr, int ix)
{
    adiag += (size_t)blockIdx.x * 3 * cs;  // cs = np: constants of each realization, else shared
    if (nxr && nxr[blockIdx.x] == ix) adiag += cs;  // this realization's last slice: its last diagonal
    bvec += (size_t)blockIdx.x * cs;
    constexpr int NP = 16 * WV, NT = 64 * WV;
    __shared__ NopivShared<WV> sh;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, c = lane & 15, g = lane >> 4;
    const int b = blockIdx.x;
    float2 *G = gss + (size_t)b * NP * NP;

    if (cqv.cq) rg_load_cq(sh.cq, cqv, b, tid, NT);
    for (int i = tid; i < NP; i += NT) {
        sh.bv[i] = i < n ? bvec[i] : cmk(0.0, 0.0);
        sh.ad[i] = i < n ? adiag[i] : cmk(0.0, 0.0);
    }
    if (tid == 0) sh.bad = 0;
    __syncthreads();

    // M = (diag(a) - V) - B Gss B in FP64, rounded into the accumulators; identity on the padding.
    np_f4 ar[WV], ai[WV];
    const int col = 16 * w + c;
#pragma unroll
    for (int ti = 0; ti < WV; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * ti + 4 * g + v;
            double2 m;
            if (row < n && col < n) {
                m = row == col ? sh.ad[row] : cmk(0.0, 0.0);
                if (cqv.cq) {
                    if (boundtype == RG_DIRICHLET) {
                        m.x -= sh.cq[abs(row - col)] - sh.cq[row + col + 2];
                    } else {
                        const int dq = col - row;
                        const double2 dd = dq >= 0 ? cmk(sh.cq[2 * dq], sh.cq[2 * dq + 1])
                                                   : cmk(sh.cq[-2 * dq], -sh.cq[-2 * dq + 1]);
                        m = csub(m, dd);
                    }
                }
                if (!first) {
                    const float2 gv = G[(size_t)row * NP + col];
                    m = csub(m, cmul(cmul(sh.bv[row], cmk((double)gv.x, (double)gv.y)), sh.bv[col]));
                }
            } else {
                m = cmk(row == col ? 1.0 : 0.0, 0.0);
            }
            ar[ti][v] = (float)m.x;
            ai[ti][v] = (float)m.y;
        }

    // Accumulator slot s holds row block (s + K) mod WV during block step K: the block being
    // eliminated is always slot 0, the next one slot 1 (all register indices static).
    if (w == 0) {
        publish_inverse<WV, 0>(ar, ai, 0, lane, sh);
        publish_column<WV, 0>(ar, ai, 0, 0, lane, sh);
    }
    __syncthreads();
    for (int K = 0; K < WV; K++) {
        const int buf = K & 1;
        // Rt = D^-1 X[K, w]; wavefront K: Rt = D^-1 and its other tiles restart from zero
        np_f4 rr, ri;
        if (w == K) {
            load_b(rr, ri, sh.dtr[buf], sh.dti[buf], lane);
#pragma unroll
            for (int s = 1; s < WV; s++) {
                ar[s] = np_f4{0.f, 0.f, 0.f, 0.f};
                ai[s] = np_f4{0.f, 0.f, 0.f, 0.f};
            }
        } else {
            tile_mul(rr, ri, sh.dr[buf], sh.di[buf], ar[0], ai[0], lane);
        }
        ar[0] = rr;
        ai[0] = ri;
        // verification of the pivots of block K against the rows of block w (w > K)
        if (w > K) {
            np_f4 br, bi, vr = np_f4{0.f, 0.f, 0.f, 0.f}, vi = np_f4{0.f, 0.f, 0.f, 0.f};
            load_b(br, bi, sh.wtr[buf], sh.wti[buf], lane);
            tile_fnms(vr, vi, sh.pr[buf], sh.pi[buf], 16 * w, br, bi, lane);  // V = -P_w W
            bool ok = true;
            const double ua = sh.ua[buf][c];
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (fabs((double)vr[v]) + fabs((double)vi[v]) > NP_MARGIN * ua) ok = false;
            if (!__all(ok) && lane == 0) sh.bad = 1;
        }
        // X[i, w] -= P_i Rt for the other row blocks; the next block's owner updates its diagonal
        // tile (slot 1) first and inverts it before the rest (look-ahead)
        if (WV > 1) tile_fnms(ar[1 % WV], ai[1 % WV], sh.pr[buf], sh.pi[buf], 16 * ((1 + K) % WV), rr, ri, lane);
        __builtin_amdgcn_sched_barrier(0);
        if (w == K + 1) publish_inverse<WV, 1 % WV>(ar, ai, K + 1, lane, sh);
#pragma unroll
        for (int s = 2; s < WV; s++) {
            tile_fnms(ar[s], ai[s], sh.pr[buf], sh.pi[buf], 16 * ((s + K) % WV), rr, ri, lane);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (w == K + 1) publish_column<WV, 1 % WV>(ar, ai, K + 1, K, lane, sh);
        // rotate: slot s <- slot s + 1
        const np_f4 t0r = ar[0], t0i = ai[0];
#pragma unroll
        for (int s = 0; s + 1 < WV; s++) {
            ar[s] = ar[s + 1];
            ai[s] = ai[s + 1];
        }
        ar[WV - 1] = t0r;
        ai[WV - 1] = t0i;
        __syncthreads();
    }
    if (sh.bad) {
        if (tid == 0) redo[b] = 1;
        return;
    }
    if (tid == 0) redo[b] = 0;
    // write X = inv(M): stage 16 rows at a time through the (now free) P buffers
    float(*str)[LDW] = sh.pr[0];
    float(*sti)[LDW] = sh.pi[0];
    for (int ch = 0; ch < WV; ch++) {
        __syncthreads();
#pragma unroll
        for (int s2 = 0; s2 < WV; s2++)
            if (s2 == ch)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    str[16 * w + 4 * g + v][c] = ar[s2][v];  // rows of chunk ch, columns of block w
                    sti[16 * w + 4 * g + v][c] = ai[s2][v];
                }
        __syncthreads();
        for (int e = tid; e < 16 * NP; e += NT) {
            const int r = e / NP, j = e % NP;  // row 16 ch + r, column j
            G[(size_t)(16 * ch + r) * NP + j] = make_float2(str[16 * (j >> 4) + r][j & 15], sti[16 * (j >> 4) + r][j & 15]);
        }
    }
}

// ---- Complex symmetric slice matrices (Dirichlet guides) -------------------------------------------
//
// The complex64 counterpart of k_inverse_sym64 (rg_nopiv64.hip): M = A_x - V_x - B Gss B is complex
// symmetric on Dirichlet guides, and the block sweep operator on the lower triangle of tiles, with
// pivot block K, D^-1 = S_KK^-1 and the panel Q_I = S_IK,
//   S_IJ -= Q_I D^-1 Q_J^T (I >= J, I, J != K),  S_IK := R_I = Q_I D^-1,  S_KK := -D^-1,
// leaves S = -M^-1; the pivot checks are those of k_inverse_nopiv, in FP64.  Wave w holds the tiles
// w + NW s (rg_sym_tile) in slot s, so with NW < WV a wave holds several diagonal tiles.  Block step K:
//   A: wave w forms R_I = Q_I D^-1 into LDS for its I = w, w + NW, ... (I != K); rows of block
//      q > K are checked against V = -Q_q W;
//   B1: the owner of tile (K+1, K+1) updates it and inverts it first (D^-1, W and the pivots of the
//      next block to LDS), then every wave updates or replaces its tiles;
//   B2: the tiles in block row/column K+1 are stored as the next panel.
// One panel buffer and one D^-1 buffer, three barriers per step; the slice constants share the R
// buffer (they are read only while the slice matrix is formed).  The per-realization state is small
// (36 tiles over NW waves, 36 KB of LDS at a row stride of 16), so several realizations share a CU
// and hide each other's serial block inversions and barrier waits.  The diagonal tiles are written
// symmetrized, so every slice matrix is exactly symmetric.
template <int WV, int S>
struct Sym32Shared {
    static constexpr int NP = 16 * WV;
    alignas(16) float qr[NP][S], qi[NP][S];  // panel Q_I = S_IK [row][k] of the current block
    union {
        struct {
            alignas(16) float rr[NP][S], ri[NP][S];  // R_I = Q_I D^-1 [row][k]
        } r;
        struct {  // the slice constants, read only while the slice matrix is formed
            alignas(16) double cq[2 * NP + 8];  // 2 n + 1 Dirichlet table entries
            double2 bv[NP], ad[NP];
        } m;
    } u;
    alignas(16) float dr[16][S], di[16][S];    // D^-1 of the current block, row-major (symmetric)
    alignas(16) float wtr[16][S], wti[16][S];  // W = U_D^-1 diag(u), transposed
    double ua[16];
    int bad;
};

// The owner of the next diagonal block inverts its tile in place and publishes D^-1, W and the pivot
// magnitudes.
template <int WV, int S>
__device__ __attribute__((always_inline)) inline void publish_sym32(np_f4 &xr, np_f4 &xi, int lane,
                                                                    Sym32Shared<WV, S> &sh)
{
    const int c = lane & 15, g = lane >> 4;
    np_f4 lr, li;
    const bool ok = block_gj(xr, xi, lr, li, sh.ua, lane);
    if (!ok && lane == 0) sh.bad = 1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        sh.dr[4 * g + v][c] = xr[v];
        sh.di[4 * g + v][c] = xi[v];
    }
    __builtin_amdgcn_wave_barrier();
    np_f4 wr, wi;
    tile_mul(wr, wi, sh.dr, sh.di, lr, li, lane);  // W = D^-1 (L_D diag(u))
    *reinterpret_cast<np_f4 *>(&sh.wtr[c][4 * g]) = wr;
    *reinterpret_cast<np_f4 *>(&sh.wti[c][4 * g]) = wi;
}

// Tile (I, J) into the panel of pivot block P (the tile lies in block row or column P, off the diagonal).
template <int WV, int S>
__device__ __attribute__((always_inline)) inline void panel_store_sym32(np_f4 xr, np_f4 xi, int I, int J, int P,
                                                                        int lane, Sym32Shared<WV, S> &sh)
{
    I = __builtin_amdgcn_readfirstlane(I);  // addresses recomputed per step (see sweep_tile32)
    J = __builtin_amdgcn_readfirstlane(J);
    asm volatile("" : "+s"(I), "+s"(J));
    const int c = lane & 15, g = lane >> 4;
    if (J == P) {  // Q_I = S_IP: the tile as it stands
#pragma unroll
        for (int v = 0; v < 4; v++) {
            sh.qr[16 * I + 4 * g + v][c] = xr[v];
            sh.qi[16 * I + 4 * g + v][c] = xi[v];
        }
    } else {  // Q_J = S_PJ^T: lane (c, g) holds Q_J[c][4g + v]
        *reinterpret_cast<np_f4 *>(&sh.qr[16 * J + c][4 * g]) = xr;
        *reinterpret_cast<np_f4 *>(&sh.qi[16 * J + c][4 * g]) = xi;
    }
}

// Block step K on tile (I, J): replaced when it lies in block row or column K, else updated.
template <int WV, int S>
__device__ __attribute__((always_inline)) inline void sweep_tile32(np_f4 &xr, np_f4 &xi, int I, int J, int K,
                                                                   int lane, Sym32Shared<WV, S> &sh)
{
    // the tile's LDS addresses are recomputed at every step: hoisted out of the step loop, those of all
    // slots would stay live and spill
    I = __builtin_amdgcn_readfirstlane(I);
    J = __builtin_amdgcn_readfirstlane(J);
    asm volatile("" : "+s"(I), "+s"(J));
    const int c = lane & 15, g = lane >> 4;
    if (I == K && J == K) {  // -D^-1 (this wave inverted the tile in place)
        xr = -xr;
        xi = -xi;
    } else if (J == K) {  // R_I
#pragma unroll
        for (int v = 0; v < 4; v++) {
            xr[v] = sh.u.r.rr[16 * I + 4 * g + v][c];
            xi[v] = sh.u.r.ri[16 * I + 4 * g + v][c];
        }
    } else if (I == K) {  // R_J^T
        load_b(xr, xi, sh.u.r.rr + 16 * J, sh.u.r.ri + 16 * J, lane);
    } else {  // S_IJ -= Q_I R_J^T
        np_f4 br, bi;
        load_b(br, bi, sh.u.r.rr + 16 * J, sh.u.r.ri + 16 * J, lane);
        tile_fnms(xr, xi, sh.qr, sh.qi, 16 * I, br, bi, lane);
    }
}

template <int WV, int NW, int S, int OCC>
__global__ __launch_bounds__(64 * NW, OCC) void k_inverse_sym32(float2 *__restrict__ gss, const rg_cqv cqv,
                                                                const double2 *__restrict__ adiag,
                                                                const double2 *__restrict__ bvec, int n, int first,
                                                                int32_t *__restrict__ redo, int cs,
                                                                const int32_t *__restrict__ nxr, int ix)
{
    adiag += (size_t)blockIdx.x * 3 * cs;
    if (nxr && nxr[blockIdx.x] == ix) adiag += cs;
    bvec += (size_t)blockIdx.x * cs;
    constexpr int NP = 16 * WV, NT = 64 * NW, NTILE = WV * (WV + 1) / 2, NS = (NTILE + NW - 1) / NW;
    constexpr int ND = (WV + NW - 1) / NW;  // slots that can hold a diagonal tile
    static_assert(NW <= WV, "every wave holds a diagonal tile");
    __shared__ Sym32Shared<WV, S> sh;
    const int tid = threadIdx.x, lane = tid & 63, c = lane & 15, g = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int b = blockIdx.x;
    float2 *G = gss + (size_t)b * NP * NP;

    int sI[NS], sJ[NS];
    bool sv[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        sv[s] = w + NW * s < NTILE;
        sI[s] = sJ[s] = 0;
        if (sv[s]) rg_sym_tile(w + NW * s, WV, sI[s], sJ[s]);
    }
    // The previous Gss on the lower tiles, loaded before the constants are staged so that the loads
    // overlap it.
    np_f4 ar[NS], ai[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        ar[s] = np_f4{0.f, 0.f, 0.f, 0.f};
        ai[s] = np_f4{0.f, 0.f, 0.f, 0.f};
        if (!sv[s] || first) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float2 gv = G[(size_t)(16 * sI[s] + 4 * g + v) * NP + 16 * sJ[s] + c];
            ar[s][v] = gv.x;
            ai[s][v] = gv.y;
        }
    }
    if (cqv.cq) rg_load_cq(sh.u.m.cq, cqv, b, tid, NT);
    for (int i = tid; i < NP; i += NT) {
        sh.u.m.bv[i] = i < n ? bvec[i] : cmk(0.0, 0.0);
        sh.u.m.ad[i] = i < n ? adiag[i] : cmk(0.0, 0.0);
    }
    if (tid == 0) sh.bad = 0;
    __syncthreads();

    // M = (diag(a) - V) - B Gss B in FP64 on the lower tiles, rounded; identity on the padding.
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (!sv[s]) continue;
        const int col = 16 * sJ[s] + c;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * sI[s] + 4 * g + v;
            double2 m;
            if (row < n && col < n) {
                m = row == col ? sh.u.m.ad[row] : cmk(0.0, 0.0);
                if (cqv.cq) m.x -= sh.u.m.cq[abs(row - col)] - sh.u.m.cq[row + col + 2];
                if (!first)
                    m = csub(m, cmul(cmul(sh.u.m.bv[row], cmk((double)ar[s][v], (double)ai[s][v])), sh.u.m.bv[col]));
            } else {
                m = cmk(row == col ? 1.0 : 0.0, 0.0);
            }
            ar[s][v] = (float)m.x;
            ai[s][v] = (float)m.y;
        }
    }
    // Panel of block 0 and its inverse (the constants are not read after this point: the barrier
    // below hands their space to R).
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (sv[s] && sJ[s] == 0 && sI[s] > 0) panel_store_sym32(ar[s], ai[s], sI[s], 0, 0, lane, sh);
    if (w == 0) publish_sym32(ar[0], ai[0], lane, sh);
    __syncthreads();

    for (int K = 0; K < WV; K++) {
        // A: R_I = Q_I D^-1 for this wave's I; the rows of block q > K against the pivots of block K.
#pragma unroll
        for (int r = 0; r < ND; r++) {
            int I = w + NW * r;
            if (I >= WV || I == K) continue;
            asm volatile("" : "+s"(I));  // addresses recomputed per step (see sweep_tile32)
            np_f4 br, bi, xr, xi;
            load_b(br, bi, sh.dr, sh.di, lane);
            tile_mul(xr, xi, sh.qr + 16 * I, sh.qi + 16 * I, br, bi, lane);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                sh.u.r.rr[16 * I + 4 * g + v][c] = xr[v];
                sh.u.r.ri[16 * I + 4 * g + v][c] = xi[v];
            }
        }
        for (int q = K + 1; q < WV; q++) {
            if ((WV + q) % NW != w) continue;
            np_f4 br, bi, vr = np_f4{0.f, 0.f, 0.f, 0.f}, vi = np_f4{0.f, 0.f, 0.f, 0.f};
            load_b(br, bi, sh.wtr, sh.wti, lane);
            tile_fnms(vr, vi, sh.qr, sh.qi, 16 * q, br, bi, lane);  // V = -Q_q W
            bool ok = true;
            const double ua = sh.ua[c];
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (fabs((double)vr[v]) + fabs((double)vi[v]) > NP_MARGIN * ua) ok = false;
            if (!__all(ok) && lane == 0) sh.bad = 1;
        }
        __syncthreads();
        // B1: the next diagonal tile first, then the rest.
        if (K + 1 < WV) {
#pragma unroll
            for (int s = 0; s < ND; s++) {
                if (!sv[s] || sI[s] != K + 1 || sJ[s] != K + 1) continue;
                sweep_tile32(ar[s], ai[s], K + 1, K + 1, K, lane, sh);
                publish_sym32(ar[s], ai[s], lane, sh);
            }
        }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            if (!sv[s] || (s < ND && sI[s] == K + 1 && sJ[s] == K + 1)) continue;
            sweep_tile32(ar[s], ai[s], sI[s], sJ[s], K, lane, sh);
        }
        __syncthreads();
        // B2: the next panel.
        if (K + 1 < WV) {
#pragma unroll
            for (int s = 0; s < NS; s++)
                if (sv[s] && (sI[s] == K + 1) != (sJ[s] == K + 1))
                    panel_store_sym32(ar[s], ai[s], sI[s], sJ[s], K + 1, lane, sh);
            __syncthreads();
        }
    }
    if (sh.bad) {
        if (tid == 0) redo[b] = 1;
        return;
    }
    if (tid == 0) redo[b] = 0;
    // Gss = -S, both triangles, each diagonal tile symmetrized (X + X^T) / 2 through the (free) panel
    // buffer.
#pragma unroll
    for (int s = 0; s < ND; s++) {
        if (!sv[s] || sI[s] != sJ[s]) continue;
        const int I = sI[s];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            sh.qr[16 * I + 4 * g + v][c] = ar[s][v];
            sh.qi[16 * I + 4 * g + v][c] = ai[s][v];
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < ND; s++) {
        if (!sv[s] || sI[s] != sJ[s]) continue;
        const int I = sI[s];
        const np_f4 tr = *reinterpret_cast<const np_f4 *>(&sh.qr[16 * I + c][4 * g]);
        const np_f4 tq = *reinterpret_cast<const np_f4 *>(&sh.qi[16 * I + c][4 * g]);
        ar[s] = 0.5f * (ar[s] + tr);
        ai[s] = 0.5f * (ai[s] + tq);
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (!sv[s]) continue;
        const int I = sI[s], J = sJ[s];
#pragma unroll
        for (int v = 0; v < 4; v++)
            G[(size_t)(16 * I + 4 * g + v) * NP + 16 * J + c] = make_float2(-ar[s][v], -ai[s][v]);
        if (I != J) {  // the transpose: row 16 J + c, columns 16 I + 4 g .. + 3 (32 contiguous bytes)
            float4 *dst = reinterpret_cast<float4 *>(G + (size_t)(16 * J + c) * NP + 16 * I + 4 * g);
            dst[0] = make_float4(-ar[s][0], -ai[s][0], -ar[s][1], -ai[s][1]);
            dst[1] = make_float4(-ar[s][2], -ai[s][2], -ar[s][3], -ai[s][3]);
        }
    }
}

}  // namespace

// Complex symmetric slice matrices (Dirichlet guides), 32 <= np <= 256, complex64 state: 4 waves per
// realization up to np = 128 (36 KB of LDS at np = 128, so that four realizations share a CU), one wave
// per 16-column block above (np = 256: 16 waves, 70 KB).
void rg_launch_inverse_sym32(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_INVERSE, &stop);
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    const double2 *ad = P->d_adiag + (size_t)slot * P->np;
    switch (P->np / 16) {
// OCC: waves per SIMD the register budget is set for.  At 10 and 11 blocks one realization fills a CU at 3
// or 4 waves per SIMD alike, and 3 removes the 59 spilled registers of the 4-wave budget (1.80 against 1.86 ms
// per launch of 4096 at N = 150).  Kept at 4 where measured faster: order 112 with 59 spills (0.945 ms per
// launch of 8192 at N = 100, against 1.03 at 3 waves and 1.29 with 7 waves per realization), and 224 and 256
// with 15 and 6 (3.62 and 4.32 ms, against 4.21 and 4.97 ms with half the waves at 2 per SIMD).
#define RG_SYM(W, NW, OCC)                                                                                      \
    case W:                                                                                                     \
        hipLaunchKernelGGL((k_inverse_sym32<W, NW, 16, OCC>), dim3(nb), dim3(64 * NW), 0, P->stream, P->d_gss32, \
                           cq, ad, P->d_b, P->nmu, first, redo, P->cstride, P->d_nx_r, P->cur_ix);              \
        break;
        RG_SYM(2, 2, 4)
        RG_SYM(3, 3, 4)
        RG_SYM(4, 4, 4)
        RG_SYM(5, 4, 4)
        RG_SYM(6, 4, 4)
        RG_SYM(7, 4, 4)
        RG_SYM(8, 4, 4)
        RG_SYM(9, 9, 4)
        RG_SYM(10, 10, 3)
        RG_SYM(11, 11, 3)
        RG_SYM(12, 12, 4)
        RG_SYM(13, 13, 4)
        RG_SYM(14, 14, 4)
        RG_SYM(15, 15, 4)
        RG_SYM(16, 16, 4)
#undef RG_SYM
        default: break;
    }
    rg_timer_end(P, RGK_INVERSE, stop);
}

// np >= 32 (with a single 16-column block the pivoted kernel is used directly).
void rg_launch_inverse_nopiv(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_INVERSE, &stop);
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    const double2 *ad = P->d_adiag + (size_t)slot * P->np;
    switch (P->np / 16) {
#define RG_NP(W)                                                                                              \
    case W:                                                                                                   \
        hipLaunchKernelGGL(k_inverse_nopiv<W>, dim3(nb), dim3(64 * W), 0, P->stream, P->d_gss32, cq,           \
                           P->p.boundtype, ad, P->d_b, P->nmu, first, redo, P->cstride, P->d_nx_r, P->cur_ix); \
        break;
        RG_NP(2)
        RG_NP(3)
        RG_NP(4)
        RG_NP(5)
        RG_NP(6)
        RG_NP(7)
        RG_NP(8)
#undef RG_NP
        default: break;
    }
    rg_timer_end(P, RGK_INVERSE, stop);
}
