// rg_nopiv64.hip -- blocked Gauss-Jordan slice inversion with verified diagonal pivots, complex128
// state on the FP64 matrix cores (gfx950).
//
// Slice inversion of the FP64 sweeps, src/waveguide.f90:434-435 (zgetrf + zgetri of
// src/base_utils.f90:235-274).  As in the complex64 kernel (rg_nopiv.hip): zgetrf's pivot on the
// sweep's slice matrices is the diagonal element, so the elimination takes the diagonal pivots in
// 16-column blocks and verifies that each is the one zgetrf's rule (max |Re|+|Im| over the rows not
// pivoted yet) picks, within a relative margin of 1e-4: inside a block from the elimination itself,
// below it from V = P U_D^-1 diag(u).  A realization that fails the check (or meets a zero pivot) is
// not written; it is flagged in `redo` and the pivoted kernel (k_gj_lookahead) inverts it again from
// the unchanged Gss.
//
// Layout: wavefront w owns the 16-column block w as np/16 accumulator tiles of v_mfma_f64_16x16x4f64
// (real and imaginary planes): register v of lane l holds row 4v + l/16, column l%16.  A k-step v of
// that instruction takes k = 4v + g from lane group g for both operands, so an accumulator tile is a
// valid B operand as it stands (register v at step v), and an A operand is read from LDS where every
// 16-wide k dimension is stored in the order kp(k) = 4 (k % 4) + k / 4: lane group g then reads the
// k = g, 4+g, 8+g, 12+g of a row as four consecutive doubles.  Block step K (D = tile (K, K), P =
// column block K):
//   wavefront K: Gauss-Jordan on D in registers (16 unrolled steps; the pivot by readlane, the pivot
//   row by ds_bpermute, the pivot column by DPP, each double as two halves; pivot checks
//   after the last step) -> D^-1, the pivots and W = D^-1 L_D diag(u) = U_D^-1 diag(u), to LDS with P;
//   every wavefront: Rt = D^-1 X[K, w], X[K, w] := Rt, X[i, w] -= P_i Rt (i != K) (wavefront K:
//   X[i, K] := -P_i D^-1, X[K, K] := D^-1), and the rows of its own block index checked against V.
// Look-ahead: wavefront K+1 updates its diagonal tile first and inverts the next D before its other
// tiles, so one workgroup barrier per block step.  The 1/|p|^2 of a pivot is v_rcp_f64 with two Newton
// steps (full FP64 accuracy).
#include "rg_internal.h"
#include <utility>

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr double N64_MARGIN = 1.0 - 1e-4;
constexpr int LDD = 20;  // LDS row stride (doubles): 32-byte aligned rows, four k values per read

__device__ __forceinline__ constexpr int kp(int k) { return 4 * (k & 3) + (k >> 2); }

template <int WV>
struct Nopiv64Shared {
    static constexpr int NP = 16 * WV;
    alignas(32) double pr[2][NP][LDD], pi[2][NP][LDD];    // P (column block, rows outside the block; zero inside)
    alignas(32) double dr[2][16][LDD], di[2][16][LDD];    // D^-1 [row][kp(col)]
    alignas(32) double dtr[2][16][LDD], dti[2][16][LDD];  // D^-1 transposed [col][kp(row)] (B operand reads)
    alignas(32) double wtr[2][16][LDD], wti[2][16][LDD];  // W = U_D^-1 diag(u), transposed
    double ua[2][16];                                     // |u_j|_1
    int bad;
    alignas(16) double cq[2 * 256 + 8];
    double2 bv[NP], ad[NP];  // b and the slice diagonal
};

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Lane (c, g) <- lane (J, g): DPP row_newbcast:J on both halves.
template <int J>
__device__ __forceinline__ double lane_bcast_d(double x)
{
    const long long b = __double_as_longlong(x);
    // every lane is written (all rows and banks), so no old value is needed
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x150 + J, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x150 + J, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double rcp_d(double q)
{
    double r = __builtin_amdgcn_rcp(q);
    r = fma(fma(-q, r, 1.0), r, r);
    r = fma(fma(-q, r, 1.0), r, r);
    return r;
}

// Step J of the Gauss-Jordan elimination of a 16x16 block (xr, xi) held in accumulator layout by one
// wavefront, diagonal pivot.  Column J is reset to e_J, the scaled pivot row rv = a[J,:]/p (rv_J = 1/p)
// is formed, and every row takes one uniform update x -= m rv with m = a[i,J], and m = p - 1 for the
// pivot row (which turns it into a[J,:]/p).  Lane c == J records |p| and the reduced column (rows >= J)
// for the checks and for L_D diag(u).
template <int J>
__device__ __forceinline__ void gj_step64(d4 &xr, d4 &xi, d4 &lr, d4 &li, double &my_ua, bool &my_nz, int lane)
{
    constexpr int VJ = J >> 2, G = J & 3;
    const int c = lane & 15, g = lane >> 4;
    const double pvr = readlane_d(xr[VJ], J + 16 * G), pvi = readlane_d(xi[VJ], J + 16 * G);
    const double prr = __shfl(xr[VJ], c + 16 * G, 64), pri = __shfl(xi[VJ], c + 16 * G, 64);  // the pivot row
    double mr[4], mi[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        mr[v] = lane_bcast_d<J>(xr[v]);
        mi[v] = lane_bcast_d<J>(xi[v]);
    }
    const double den = pvr * pvr + pvi * pvi;
    const double inv = den > 0.0 ? rcp_d(den) : 1.0;
    const double dr = pvr * inv, di = -pvi * inv;
    if (c == J) {
        my_ua = fabs(pvr) + fabs(pvi);
        my_nz = den > 0.0;
    }
    if (c == J) {  // lane J takes its column before the update (moves under the lane mask)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            lr[v] = mr[v];
            li[v] = mi[v];
        }
    }
    asm volatile("" : "+v"(lr), "+v"(li));  // now: else the 16 steps' columns all stay live
    double rvr, rvi;
    if (c == J) {
        rvr = dr;
        rvi = di;
    } else {
        rvr = prr * dr - pri * di;
        rvi = prr * di + pri * dr;
    }
    // x <- keep x - m rv (+ 1 at (J, J)), keep = 0 on column J: the column's reset to e_J and the update
    // in one FMA per plane
    const double keep = c == J ? 0.0 : 1.0;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        double mur = mr[v], mui = mi[v];
        if (v == VJ && g == G) {  // row J
            mur = pvr - 1.0;
            mui = pvi;
        }
        double nr = fma(mui, rvi, -(mur * rvr));  // -Re(m rv)
        const double ni = fma(-mur, rvi, -(mui * rvr));  // -Im(m rv)
        if (v == VJ) nr += (g == G && c == J) ? 1.0 : 0.0;
        xr[v] = fma(keep, xr[v], nr);
        xi[v] = fma(keep, xi[v], ni);
    }
}

template <int... Js>
__device__ __forceinline__ void gj_steps64(std::integer_sequence<int, Js...>, d4 &xr, d4 &xi, d4 &lr, d4 &li,
                                           double &my_ua, bool &my_nz, int lane)
{
    (gj_step64<Js>(xr, xi, lr, li, my_ua, my_nz, lane), ...);
}

// Gauss-Jordan on the 16x16 block, 16 steps unrolled.  On return (xr, xi) = D^-1, (lr, li) =
// L_D diag(u), ua[j] = |u_j|_1; false if a pivot fails the zgetrf check or is zero.
__device__ __forceinline__ bool block_gj64(d4 &xr, d4 &xi, d4 &lr, d4 &li, double *ua, int lane)
{
    const int c = lane & 15, g = lane >> 4;
    lr = d4{0.0, 0.0, 0.0, 0.0};
    li = d4{0.0, 0.0, 0.0, 0.0};
    double my_ua = 0.0;
    bool my_nz = true;
    int ln = lane;
    asm volatile("" : "+v"(ln));  // keep the per-step lane masks inside the block loop
    gj_steps64(std::make_integer_sequence<int, 16>{}, xr, xi, lr, li, my_ua, my_nz, ln);
    bool ok = my_ua > 0.0 && my_nz;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int row = 4 * v + g;
        if (row < c) {  // the capture took the whole column: L_D has no entries above the diagonal
            lr[v] = 0.0;
            li[v] = 0.0;
        }
        const bool big = fabs(lr[v]) + fabs(li[v]) > N64_MARGIN * my_ua;
        ok = ok && !(row > c && big);
    }
    if (lane < 16) ua[c] = my_ua;
    return __all(ok);
}

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// C (tile) -= A B for one 16x16 complex tile: A rows r0.. from LDS ([row][kp(k)]), B in accumulator layout.
__device__ __forceinline__ void tile_fnms64(d4 &cr, d4 &ci, const double (*asr)[LDD], const double (*asi)[LDD], int r0,
                                            d4 br, d4 bi, int lane)
{
    const int i = lane & 15, g = lane >> 4;
    const d4 a_r = *reinterpret_cast<const d4 *>(&asr[r0 + i][4 * g]);
    const d4 a_i = *reinterpret_cast<const d4 *>(&asi[r0 + i][4 * g]);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        cr = mfma64(-a_r[v], br[v], cr);
        cr = mfma64(a_i[v], bi[v], cr);
        ci = mfma64(-a_r[v], bi[v], ci);
        ci = mfma64(-a_i[v], br[v], ci);
    }
}

// C = A B for A from LDS [16][LDD] and B in accumulator layout.
__device__ __forceinline__ void tile_mul64(d4 &cr, d4 &ci, const double (*asr)[LDD], const double (*asi)[LDD], d4 br,
                                           d4 bi, int lane)
{
    const int i = lane & 15, g = lane >> 4;
    const d4 a_r = *reinterpret_cast<const d4 *>(&asr[i][4 * g]);
    const d4 a_i = *reinterpret_cast<const d4 *>(&asi[i][4 * g]);
    cr = d4{0.0, 0.0, 0.0, 0.0};
    ci = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        cr = mfma64(a_r[v], br[v], cr);
        cr = mfma64(-a_i[v], bi[v], cr);
        ci = mfma64(a_r[v], bi[v], ci);
        ci = mfma64(a_i[v], br[v], ci);
    }
}

// B operand (accumulator layout) from a transposed LDS matrix St[col][kp(row)]: one read per plane.
__device__ __forceinline__ void load_b64(d4 &br, d4 &bi, const double (*tr)[LDD], const double (*ti)[LDD], int lane)
{
    const int c = lane & 15, g = lane >> 4;
    br = *reinterpret_cast<const d4 *>(&tr[c][4 * g]);
    bi = *reinterpret_cast<const d4 *>(&ti[c][4 * g]);
}

// The owner of block K inverts its diagonal tile (accumulator slot SLOT) and publishes D^-1, W and the
// pivot magnitudes to buffer K & 1.
template <int WV, int SLOT>
__device__ __forceinline__ void publish_inverse64(const d4 (&ar)[WV], const d4 (&ai)[WV], int K, int lane,
                                                  Nopiv64Shared<WV> &sh)
{
    const int c = lane & 15, g = lane >> 4, buf = K & 1;
    d4 dr = ar[SLOT], di = ai[SLOT], lr, li;
    const bool ok = block_gj64(dr, di, lr, li, sh.ua[buf], lane);
    if (!ok && lane == 0) sh.bad = 1;
#pragma unroll
    for (int v = 0; v < 4; v++) {  // D^-1[4v + g][c]
        sh.dr[buf][4 * v + g][kp(c)] = dr[v];
        sh.di[buf][4 * v + g][kp(c)] = di[v];
    }
    *reinterpret_cast<d4 *>(&sh.dtr[buf][c][4 * g]) = dr;  // [c][kp(4v + g)] = [c][4g + v]
    *reinterpret_cast<d4 *>(&sh.dti[buf][c][4 * g]) = di;
    __builtin_amdgcn_wave_barrier();
    d4 wr, wi;
    tile_mul64(wr, wi, sh.dr[buf], sh.di[buf], lr, li, lane);  // W = D^-1 (L_D diag(u))
    *reinterpret_cast<d4 *>(&sh.wtr[buf][c][4 * g]) = wr;
    *reinterpret_cast<d4 *>(&sh.wti[buf][c][4 * g]) = wi;
}

// ... and the column block P of block K (rows outside the block; zero inside) to buffer K & 1.
// Slot s holds row block (s + rot) mod WV.
template <int WV, int SLOT>
__device__ __forceinline__ void publish_column64(const d4 (&ar)[WV], const d4 (&ai)[WV], int K, int rot, int lane,
                                                 Nopiv64Shared<WV> &sh)
{
    const int c = lane & 15, g = lane >> 4, buf = K & 1;
#pragma unroll
    for (int s = 0; s < WV; s++) {
        const int rb = (s + rot) % WV;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            sh.pr[buf][16 * rb + 4 * v + g][kp(c)] = s == SLOT ? 0.0 : ar[s][v];
            sh.pi[buf][16 * rb + 4 * v + g][kp(c)] = s == SLOT ? 0.0 : ai[s][v];
        }
    }
}

template <int WV>
__global__ __launch_bounds__(64 * WV, 2) void k_inverse_nopiv64(double2 *__restrict__ gss, const rg_cqv cqv,
                                                                int boundtype, const double2 *__restrict__ adiag,
                                                                const double2 *__restrict__ bvec, int n, int first,
                                                                int32_t *__restrict__ redo, int cs,
                                                                const int32_t *__restrict__ nxr, int ix)
{
    adiag += (size_t)blockIdx.x * 3 * cs;  // cs = np: constants of each realization, else shared
    if (nxr && nxr[blockIdx.x] == ix) adiag += cs;  // this realization's last slice: its last diagonal
    bvec += (size_t)blockIdx.x * cs;
    constexpr int NP = 16 * WV, NT = 64 * WV;
    __shared__ Nopiv64Shared<WV> sh;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, c = lane & 15, g = lane >> 4;
    const int b = blockIdx.x;
    double2 *G = gss + (size_t)b * NP * NP;

    if (cqv.cq) rg_load_cq(sh.cq, cqv, b, tid, NT);
    for (int i = tid; i < NP; i += NT) {
        sh.bv[i] = i < n ? bvec[i] : cmk(0.0, 0.0);
        sh.ad[i] = i < n ? adiag[i] : cmk(0.0, 0.0);
    }
    if (tid == 0) sh.bad = 0;
    __syncthreads();

    // M = (diag(a) - V) - B Gss B into the accumulators; identity on the padding.
    d4 ar[WV], ai[WV];
    const int col = 16 * w + c;
#pragma unroll
    for (int ti = 0; ti < WV; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * ti + 4 * v + g;
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
                if (!first) m = csub(m, cmul(cmul(sh.bv[row], G[(size_t)row * NP + col]), sh.bv[col]));
            } else {
                m = cmk(row == col ? 1.0 : 0.0, 0.0);
            }
            ar[ti][v] = m.x;
            ai[ti][v] = m.y;
        }

    // Slot s holds row block (s + K) mod WV during block step K: the block being eliminated is
    // always slot 0, the next one slot 1 (all register indices static).
    if (w == 0) {
        publish_inverse64<WV, 0>(ar, ai, 0, lane, sh);
        publish_column64<WV, 0>(ar, ai, 0, 0, lane, sh);
    }
    __syncthreads();
    for (int K = 0; K < WV; K++) {
        const int buf = K & 1;
        d4 rr, ri;
        if (w == K) {
            load_b64(rr, ri, sh.dtr[buf], sh.dti[buf], lane);
#pragma unroll
            for (int s = 1; s < WV; s++) {
                ar[s] = d4{0.0, 0.0, 0.0, 0.0};
                ai[s] = d4{0.0, 0.0, 0.0, 0.0};
            }
        } else {
            tile_mul64(rr, ri, sh.dr[buf], sh.di[buf], ar[0], ai[0], lane);
        }
        ar[0] = rr;
        ai[0] = ri;
        if (w > K) {  // pivots of block K against the rows of block w
            d4 br, bi, vr = d4{0.0, 0.0, 0.0, 0.0}, vi = d4{0.0, 0.0, 0.0, 0.0};
            load_b64(br, bi, sh.wtr[buf], sh.wti[buf], lane);
            tile_fnms64(vr, vi, sh.pr[buf], sh.pi[buf], 16 * w, br, bi, lane);  // V = -P_w W
            bool ok = true;
            const double ua = sh.ua[buf][c];
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (fabs(vr[v]) + fabs(vi[v]) > N64_MARGIN * ua) ok = false;
            if (!__all(ok) && lane == 0) sh.bad = 1;
        }
        if (WV > 1) tile_fnms64(ar[1 % WV], ai[1 % WV], sh.pr[buf], sh.pi[buf], 16 * ((1 + K) % WV), rr, ri, lane);
        if (w == K + 1) publish_inverse64<WV, 1 % WV>(ar, ai, K + 1, lane, sh);
#pragma unroll
        for (int s = 2; s < WV; s++) {
            tile_fnms64(ar[s], ai[s], sh.pr[buf], sh.pi[buf], 16 * ((s + K) % WV), rr, ri, lane);
        }
        if (w == K + 1) publish_column64<WV, 1 % WV>(ar, ai, K + 1, K, lane, sh);
        const d4 t0r = ar[0], t0i = ai[0];
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
    // write X = inv(M) straight from the tiles (slot s holds row block s again after WV rotations): one
    // store writes four rows of 16 consecutive complex128 values
#pragma unroll
    for (int ti = 0; ti < WV; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) G[(size_t)(16 * ti + 4 * v + g) * NP + col] = make_double2(ar[ti][v], ai[ti][v]);
}

// ---- Complex symmetric slice matrices (Dirichlet guides) -------------------------------------------
//
// With Dirichlet modes V_x is real symmetric and A_x, B are diagonal, so M = A_x - V_x - B Gss B is
// complex symmetric whenever Gss is, and Gss = M^-1 is: every slice matrix of the sweep is.  The
// elimination then runs as the block sweep operator on the lower triangle of tiles only: for pivot
// block K with D^-1 = S_KK^-1 and the panel Q_I = S_IK,
//   S_IJ -= Q_I D^-1 Q_J^T (I >= J, I, J != K),  S_IK := R_I = Q_I D^-1,  S_KK := -D^-1,
// which after every block leaves S = -M^-1.  These are the diagonal-pivot Gauss-Jordan steps of
// k_inverse_nopiv64 with the signs of the eliminated column flipped, on WV (WV + 1) / 2 tiles instead
// of WV^2; the pivot checks are the same.  Wave w holds the tiles t = w + NW s (rg_sym_tile: diagonal
// tiles first, then the off-diagonal ones by distance) in register slot s.  Block step K:
//   A: wave w forms R_I = Q_I D^-1 into LDS for its I = w, w + NW, ... (I != K); rows of block q > K
//      are checked against V = -Q_q W;
//   B1: the owner of tile (K+1, K+1) updates it and inverts it first (D^-1, W and the pivots of the
//      next block to LDS), then every wave updates or replaces its tiles;
//   B2: the tiles in block row/column K+1 are stored as the next panel.
// One panel buffer and one D^-1 buffer, three barriers per step; the slice constants share the R
// buffer.  At np = 128 with 4 waves of 9 tiles and a row stride of 16 doubles (72 KB of LDS) two
// realizations share a CU and hide each other's serial block inversions and barrier waits.  With that
// stride every row starts in the same bank, so the 32-byte chunks of a row are XOR-swizzled by the
// row (sw).  Only the lower triangle of the previous Gss is read; both triangles of the new one are
// written, each diagonal tile symmetrized.

// Column `col` (already in kp order) of row `row`: the 4-double chunk index XOR the row's low bits.
__device__ __forceinline__ constexpr int sw(int row, int col) { return col ^ (4 * (row & 3)); }

// C (tile) -= A B with A rows r0.. from a swizzled LDS matrix ([row][sw(kp(k))]), B in accumulator layout.
template <int S>
__device__ __forceinline__ void tile_fnms64s(d4 &cr, d4 &ci, const double (*asr)[S], const double (*asi)[S], int r0,
                                             d4 br, d4 bi, int lane)
{
    const int i = lane & 15, g = lane >> 4;
    const d4 a_r = *reinterpret_cast<const d4 *>(&asr[r0 + i][sw(i, 4 * g)]);
    const d4 a_i = *reinterpret_cast<const d4 *>(&asi[r0 + i][sw(i, 4 * g)]);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        cr = mfma64(-a_r[v], br[v], cr);
        cr = mfma64(a_i[v], bi[v], cr);
        ci = mfma64(-a_r[v], bi[v], ci);
        ci = mfma64(-a_i[v], br[v], ci);
    }
}

// C = A B, A rows r0.. from a swizzled LDS matrix.
template <int S>
__device__ __forceinline__ void tile_mul64s(d4 &cr, d4 &ci, const double (*asr)[S], const double (*asi)[S], int r0,
                                            d4 br, d4 bi, int lane)
{
    const int i = lane & 15, g = lane >> 4;
    const d4 a_r = *reinterpret_cast<const d4 *>(&asr[r0 + i][sw(i, 4 * g)]);
    const d4 a_i = *reinterpret_cast<const d4 *>(&asi[r0 + i][sw(i, 4 * g)]);
    cr = d4{0.0, 0.0, 0.0, 0.0};
    ci = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        cr = mfma64(a_r[v], br[v], cr);
        cr = mfma64(-a_i[v], bi[v], cr);
        ci = mfma64(a_r[v], bi[v], ci);
        ci = mfma64(a_i[v], br[v], ci);
    }
}

// B operand (accumulator layout: register v = B[4v + g][c]) from the swizzled LDS rows r0 + c holding
// B^T, i.e. [r0 + c][sw(kp(k))] = B[k][c]: one 32-byte read per plane.
template <int S>
__device__ __forceinline__ void load_b64s(d4 &br, d4 &bi, const double (*tr)[S], const double (*ti)[S], int r0,
                                          int lane)
{
    const int c = lane & 15, g = lane >> 4;
    br = *reinterpret_cast<const d4 *>(&tr[r0 + c][sw(c, 4 * g)]);
    bi = *reinterpret_cast<const d4 *>(&ti[r0 + c][sw(c, 4 * g)]);
}

template <int WV, int S>
struct Sym64Shared {
    static constexpr int NP = 16 * WV;
    alignas(32) double qr[NP][S], qi[NP][S];  // panel Q_I = S_IK [row][sw(kp(col))] of the current block
    union {
        struct {
            alignas(32) double rr[NP][S], ri[NP][S];  // R_I = Q_I D^-1 [row][sw(kp(col))]
        } r;
        struct {  // the slice constants, read only while the slice matrix is formed
            alignas(16) double cq[2 * 128 + 8];  // np <= 128: 2 n + 1 Dirichlet table entries
            double2 bv[NP], ad[NP];
        } m;
    } u;
    alignas(32) double dr[16][S], di[16][S];    // D^-1 of the current block [row][sw(kp(col))] (symmetric)
    alignas(32) double wtr[16][S], wti[16][S];  // W = U_D^-1 diag(u), transposed
    double ua[16];
    int bad;
};

// The owner of the next diagonal block inverts its tile in place and publishes D^-1, W and the pivot
// magnitudes.
template <int WV, int S>
__device__ __forceinline__ void publish_sym64(d4 &xr, d4 &xi, int lane, Sym64Shared<WV, S> &sh)
{
    const int c = lane & 15, g = lane >> 4;
    d4 lr, li;
    const bool ok = block_gj64(xr, xi, lr, li, sh.ua, lane);
    if (!ok && lane == 0) sh.bad = 1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        sh.dr[4 * v + g][sw(4 * v + g, kp(c))] = xr[v];
        sh.di[4 * v + g][sw(4 * v + g, kp(c))] = xi[v];
    }
    __builtin_amdgcn_wave_barrier();
    d4 wr, wi;
    tile_mul64s(wr, wi, sh.dr, sh.di, 0, lr, li, lane);  // W = D^-1 (L_D diag(u))
    *reinterpret_cast<d4 *>(&sh.wtr[c][sw(c, 4 * g)]) = wr;  // [c][kp(4v + g)] = [c][4g + v]
    *reinterpret_cast<d4 *>(&sh.wti[c][sw(c, 4 * g)]) = wi;
}

// Tile (I, J) into the panel of pivot block P (the tile lies in block row or column P, off the diagonal).
template <int WV, int S>
__device__ __forceinline__ void panel_store_sym64(d4 xr, d4 xi, int I, int J, int P, int lane, Sym64Shared<WV, S> &sh)
{
    I = __builtin_amdgcn_readfirstlane(I);  // addresses recomputed per step (see sweep_tile64)
    J = __builtin_amdgcn_readfirstlane(J);
    asm volatile("" : "+s"(I), "+s"(J));
    const int c = lane & 15, g = lane >> 4;
    if (J == P) {  // Q_I = S_IP: the tile as it stands
#pragma unroll
        for (int v = 0; v < 4; v++) {
            sh.qr[16 * I + 4 * v + g][sw(4 * v + g, kp(c))] = xr[v];
            sh.qi[16 * I + 4 * v + g][sw(4 * v + g, kp(c))] = xi[v];
        }
    } else {  // Q_J = S_PJ^T: lane (c, g) holds Q_J[c][4v + g], stored at [c][kp(4v + g)] = [c][4g + v]
        *reinterpret_cast<d4 *>(&sh.qr[16 * J + c][sw(c, 4 * g)]) = xr;
        *reinterpret_cast<d4 *>(&sh.qi[16 * J + c][sw(c, 4 * g)]) = xi;
    }
}

// Block step K on tile (I, J): replaced when it lies in block row or column K, else updated.
template <int WV, int S>
__device__ __forceinline__ void sweep_tile64(d4 &xr, d4 &xi, int I, int J, int K, int lane, Sym64Shared<WV, S> &sh)
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
            xr[v] = sh.u.r.rr[16 * I + 4 * v + g][sw(4 * v + g, kp(c))];
            xi[v] = sh.u.r.ri[16 * I + 4 * v + g][sw(4 * v + g, kp(c))];
        }
    } else if (I == K) {  // R_J^T
        load_b64s(xr, xi, sh.u.r.rr, sh.u.r.ri, 16 * J, lane);
    } else {  // S_IJ -= Q_I R_J^T
        d4 br, bi;
        load_b64s(br, bi, sh.u.r.rr, sh.u.r.ri, 16 * J, lane);
        tile_fnms64s(xr, xi, sh.qr, sh.qi, 16 * I, br, bi, lane);
    }
}

template <int WV, int NW, int S, int OCC>
__global__ __launch_bounds__(64 * NW, OCC) void k_inverse_sym64(double2 *__restrict__ gss, const rg_cqv cqv,
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
    __shared__ Sym64Shared<WV, S> sh;
    const int tid = threadIdx.x, lane = tid & 63, c = lane & 15, g = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int b = blockIdx.x;
    double2 *G = gss + (size_t)b * NP * NP;

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
    d4 ar[NS], ai[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        ar[s] = d4{0.0, 0.0, 0.0, 0.0};
        ai[s] = d4{0.0, 0.0, 0.0, 0.0};
        if (!sv[s] || first) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double2 gv = G[(size_t)(16 * sI[s] + 4 * v + g) * NP + 16 * sJ[s] + c];
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

    // M = (diag(a) - V) - B Gss B on the lower tiles; identity on the padding.
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (!sv[s]) continue;
        const int col = 16 * sJ[s] + c;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * sI[s] + 4 * v + g;
            double2 m;
            if (row < n && col < n) {
                m = row == col ? sh.u.m.ad[row] : cmk(0.0, 0.0);
                if (cqv.cq) m.x -= sh.u.m.cq[abs(row - col)] - sh.u.m.cq[row + col + 2];
                if (!first) m = csub(m, cmul(cmul(sh.u.m.bv[row], cmk(ar[s][v], ai[s][v])), sh.u.m.bv[col]));
            } else {
                m = cmk(row == col ? 1.0 : 0.0, 0.0);
            }
            ar[s][v] = m.x;
            ai[s][v] = m.y;
        }
    }
    // Panel of block 0 and its inverse (the constants are not read after this point: the barrier
    // below hands their space to R).
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (sv[s] && sJ[s] == 0 && sI[s] > 0) panel_store_sym64(ar[s], ai[s], sI[s], 0, 0, lane, sh);
    if (w == 0) publish_sym64(ar[0], ai[0], lane, sh);
    __syncthreads();

    for (int K = 0; K < WV; K++) {
        // A: R_I = Q_I D^-1 for this wave's I; the rows of block q > K against the pivots of block K.
#pragma unroll
        for (int r = 0; r < ND; r++) {
            int I = w + NW * r;
            if (I >= WV || I == K) continue;
            asm volatile("" : "+s"(I));  // addresses recomputed per step (see sweep_tile64)
            d4 br, bi, xr, xi;
            load_b64s(br, bi, sh.dr, sh.di, 0, lane);  // D^-1 is symmetric: its rows serve as B^T
            tile_mul64s(xr, xi, sh.qr, sh.qi, 16 * I, br, bi, lane);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                sh.u.r.rr[16 * I + 4 * v + g][sw(4 * v + g, kp(c))] = xr[v];
                sh.u.r.ri[16 * I + 4 * v + g][sw(4 * v + g, kp(c))] = xi[v];
            }
        }
        for (int q = K + 1; q < WV; q++) {
            if ((WV + q) % NW != w) continue;
            d4 br, bi, vr = d4{0.0, 0.0, 0.0, 0.0}, vi = d4{0.0, 0.0, 0.0, 0.0};
            load_b64s(br, bi, sh.wtr, sh.wti, 0, lane);
            tile_fnms64s(vr, vi, sh.qr, sh.qi, 16 * q, br, bi, lane);  // V = -Q_q W
            bool ok = true;
            const double ua = sh.ua[c];
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (fabs(vr[v]) + fabs(vi[v]) > N64_MARGIN * ua) ok = false;
            if (!__all(ok) && lane == 0) sh.bad = 1;
        }
        __syncthreads();
        // B1: the next diagonal tile first, then the rest.
        if (K + 1 < WV) {
#pragma unroll
            for (int s = 0; s < ND; s++) {
                if (!sv[s] || sI[s] != K + 1 || sJ[s] != K + 1) continue;
                sweep_tile64(ar[s], ai[s], K + 1, K + 1, K, lane, sh);
                publish_sym64(ar[s], ai[s], lane, sh);
            }
        }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            if (!sv[s] || (s < ND && sI[s] == K + 1 && sJ[s] == K + 1)) continue;
            sweep_tile64(ar[s], ai[s], sI[s], sJ[s], K, lane, sh);
        }
        __syncthreads();
        // B2: the next panel.
        if (K + 1 < WV) {
#pragma unroll
            for (int s = 0; s < NS; s++)
                if (sv[s] && (sI[s] == K + 1) != (sJ[s] == K + 1))
                    panel_store_sym64(ar[s], ai[s], sI[s], sJ[s], K + 1, lane, sh);
            __syncthreads();
        }
    }
    if (sh.bad) {
        if (tid == 0) redo[b] = 1;
        return;
    }
    if (tid == 0) redo[b] = 0;
    // Gss = -S, both triangles, with each diagonal tile symmetrized, (X + X^T) / 2: the next slice reads
    // the diagonal tiles whole, and the rounding asymmetry of a diagonal tile is not damped by the
    // recursion (on some guides it doubles from slice to slice); written exactly symmetric, every slice
    // matrix is.  The panel buffer is free by now.
#pragma unroll
    for (int s = 0; s < ND; s++) {
        if (!sv[s] || sI[s] != sJ[s]) continue;
        const int I = sI[s];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            sh.qr[16 * I + 4 * v + g][c] = ar[s][v];
            sh.qi[16 * I + 4 * v + g][c] = ai[s][v];
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < ND; s++) {
        if (!sv[s] || sI[s] != sJ[s]) continue;
        const int I = sI[s];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            ar[s][v] = 0.5 * (ar[s][v] + sh.qr[16 * I + c][4 * v + g]);
            ai[s][v] = 0.5 * (ai[s][v] + sh.qi[16 * I + c][4 * v + g]);
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (!sv[s]) continue;
        const int I = sI[s], J = sJ[s];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double2 x = make_double2(-ar[s][v], -ai[s][v]);
            G[(size_t)(16 * I + 4 * v + g) * NP + 16 * J + c] = x;
            if (I != J) G[(size_t)(16 * J + c) * NP + 16 * I + 4 * v + g] = x;
        }
    }
}

}  // namespace

// Complex symmetric slice matrices (Dirichlet guides), np >= 32: 4 waves per realization (NW = np / 16
// below 64) and 72 KB of LDS at np = 128, so that two realizations share a CU.
void rg_launch_inverse_sym64(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_INVERSE, &stop);
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    const double2 *ad = P->d_adiag + (size_t)slot * P->np;
    switch (P->np / 16) {
#define RG_SYM(W, NW)                                                                                         \
    case W:                                                                                                   \
        hipLaunchKernelGGL((k_inverse_sym64<W, NW, 16, 2>), dim3(nb), dim3(64 * NW), 0, P->stream, P->d_gss, cq, \
                           ad, P->d_b, P->nmu, first, redo, P->cstride, P->d_nx_r, P->cur_ix);               \
        break;
        RG_SYM(2, 2)
        RG_SYM(3, 3)
        RG_SYM(4, 4)
        RG_SYM(5, 4)
        RG_SYM(6, 4)
        RG_SYM(7, 4)
        RG_SYM(8, 4)
#undef RG_SYM
        default: break;
    }
    rg_timer_end(P, RGK_INVERSE, stop);
}

// np >= 32 (with a single 16-column block the pivoted kernel is used directly).
void rg_launch_inverse_nopiv64(rg_plan *P, int nb, int slot, int first, bool with_disorder, int32_t *redo)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, RGK_INVERSE, &stop);
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    const double2 *ad = P->d_adiag + (size_t)slot * P->np;
    switch (P->np / 16) {
#define RG_NP(W)                                                                                                \
    case W:                                                                                                     \
        hipLaunchKernelGGL(k_inverse_nopiv64<W>, dim3(nb), dim3(64 * W), 0, P->stream, P->d_gss, cq,             \
                           P->p.boundtype, ad, P->d_b, P->nmu, first, redo, P->cstride, P->d_nx_r, P->cur_ix);   \
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
