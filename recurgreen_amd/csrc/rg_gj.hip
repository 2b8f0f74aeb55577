// rg_gj.hip -- register-resident Gauss-Jordan slice inversion with look-ahead pivoting (gfx950).
//
// Slice inversion of src/waveguide.f90:434-435 (zgetrf + zgetri, src/base_utils.f90:235-274):
// Gss <- inv(M), M = (diag(a_x) - V_x) - B Gss B, for every realization of the batch.
//
// The np x np matrix (np = 16*TR) is spread over 4*np threads: thread (r, c) = (tid % 16, tid / 16)
// holds rows {r + 16 i, i < TR} x columns [4c, 4c+4), so the 16 owners of a column are one DPP row
// of one wavefront.  Elimination step k is one uniform rank-1 update a <- a - m rv^T: column k is
// replaced by e_p, rv = a[p,:]/pv (so rv_k = 1/pv), and the pivot row's multiplier is pv - 1, which
// turns a[p,:] into a[p,:]/pv.  Pivots as zgetrf: max |Re|+|Im| over the rows not pivoted yet,
// lowest row on ties; rows are never swapped and the result is written as
// inv(M)[i][p_l] = X[p_i][l].
//
// Look-ahead: during step k the 16 owners of column k+1 update that column first, pick pivot
// p_{k+1}, publish its multipliers and reciprocal, reset the column to e_p, and only then update
// their other columns; so the pivot search is off the critical path and a step costs the update,
// one barrier, the broadcast of the scaled pivot row by its 32 owners, and a second barrier.
//
// T = double (complex128 sweep) or float (mixed sweeps: elimination in FP32, pivot magnitudes,
// comparisons and reciprocal pivots in FP64).  The float tile is 64 VGPRs, so two workgroups
// share a compute unit and hide each other's barrier latency.
#include "rg_internal.h"

#include <type_traits>

namespace {

template <class T> struct CT;
template <> struct CT<double> { typedef double2 C; };
template <> struct CT<float> { typedef float2 C; };

__device__ inline double2 d2(double2 v) { return v; }
__device__ inline double2 d2(float2 v) { return cmk((double)v.x, (double)v.y); }
template <class C> __device__ inline C c_from(double2 v);
template <> __device__ inline double2 c_from<double2>(double2 v) { return v; }
template <> __device__ inline float2 c_from<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ inline double2 ldc2(const double2 *p) { return *p; }
__device__ inline double2 ldc2(const float2 *p) { const float2 v = *p; return cmk((double)v.x, (double)v.y); }

template <class C>
__device__ inline C c_fnms(C acc, C m, C r)  // acc - m*r
{
    acc.x = acc.x - m.x * r.x + m.y * r.y;
    acc.y = acc.y - m.x * r.y - m.y * r.x;
    return acc;
}

template <int CTRL>
__device__ inline int dpp_row(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false); }

template <int CTRL>
__device__ __attribute__((always_inline)) inline void argmax_dpp(double &best, int &bidx)
{
    const long long bits = __double_as_longlong(best);
    const int lo = dpp_row<CTRL>((int)(bits & 0xffffffffll)), hi = dpp_row<CTRL>((int)(bits >> 32));
    const int oi = dpp_row<CTRL>(bidx);
    const double ob = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
}

// dst[0..4) <- a[ip][0..4) for a wave-uniform ip, as a chain of scalar branches (an indexed
// register access would move the tile to scratch).  The empty asm with the branch's own row
// number keeps the branches distinct, so the stores are not merged into one store from a
// run-time row index.
template <int I, int TR, class C>
__device__ __attribute__((always_inline)) inline void store_row(const C (&a)[TR][4], int ip, C *dst)
{
    if constexpr (I < TR) {
        if (ip == I) {
#pragma unroll
            for (int j = 0; j < 4; j++) dst[j] = a[I][j];
            asm volatile("; pivot row %0" ::"n"(I));
        } else {
            store_row<I + 1, TR, C>(a, ip, dst);
        }
    }
}

template <class C>
__device__ inline C c_mul(C x, C d)
{
    C y;
    y.x = x.x * d.x - x.y * d.y;
    y.y = x.x * d.y + x.y * d.x;
    return y;
}

__device__ inline double2 slice_v3(const double *s_cq, int boundtype, int m, int n)
{
    if (boundtype == RG_DIRICHLET) return cmk(s_cq[abs(m - n)] - s_cq[m + n + 2], 0.0);
    if (n >= m) return cmk(s_cq[2 * (n - m)], s_cq[2 * (n - m) + 1]);
    return cmk(s_cq[2 * (m - n)], -s_cq[2 * (m - n) + 1]);
}

// Pivot search on column jn of this thread's tile (the 16 lanes of one DPP row do it together),
// then: multipliers (pv - 1 on the pivot row) to s_col, p to s_perm[kc], 1/pv (FP64) to s_d,
// column reset to e_p.
template <int jn, int TR, class C>
__device__ __attribute__((always_inline)) inline void pivot_column(C (&a)[TR][4], const int r,
                                                                   const unsigned pivoted, const int kc, C *s_col,
                                                                   int *s_perm, double2 &s_d, int &s_sing)
{
    double best = -1.0;
    int bidx = 0x7fffffff;
#pragma unroll
    for (int i = 0; i < TR; i++) {
        if (!((pivoted >> i) & 1u)) {
            const double v = cabs1(d2(a[i][jn]));
            if (v > best) { best = v; bidx = r + 16 * i; }
        }
    }
    argmax_dpp<0xB1>(best, bidx);   // quad_perm [1,0,3,2]
    argmax_dpp<0x4E>(best, bidx);   // quad_perm [2,3,0,1]
    argmax_dpp<0x141>(best, bidx);  // row_half_mirror
    argmax_dpp<0x140>(best, bidx);  // row_mirror
#pragma unroll
    for (int i = 0; i < TR; i++) {
        const int gi = r + 16 * i;
        const bool isp = gi == bidx;
        C x = a[i][jn];
        if (isp) {
            const double2 pv = d2(x);
            const bool sing = !(best > 0.0);
            s_perm[kc] = gi;
            s_d = sing ? cmk(1.0, 0.0) : crecip(pv);
            if (sing) s_sing = 1;
            x = c_from<C>(cmk(pv.x - 1.0, pv.y));
        }
        s_col[gi] = x;
        C e;
        e.x = isp ? 1 : 0;
        e.y = 0;
        a[i][jn] = e;
    }
}

template <int TR, class T>
__global__ __launch_bounds__(64 * TR, sizeof(T) == 8 ? 2 : 4) void k_gj_lookahead(
    typename CT<T>::C *__restrict__ gss, const rg_cqv cqv, int boundtype,
    const double2 *__restrict__ adiag, const double2 *__restrict__ bvec, int n, int first, int32_t *__restrict__ status,
    const int32_t *__restrict__ only, int cs, const int32_t *__restrict__ nxr, int ix)
{
    using C = typename CT<T>::C;
    if (only && only[blockIdx.x] == 0) return;
    adiag += (size_t)blockIdx.x * 3 * cs;  // cs = np: constants of each realization, else shared
    if (nxr && nxr[blockIdx.x] == ix) adiag += cs;  // this realization's last slice: its last diagonal
    bvec += (size_t)blockIdx.x * cs;
    constexpr int NP = 16 * TR, TC = 4, NT = 64 * TR;
    constexpr int STAGE_ROWS = NP <= 16 ? 16 : (NP <= 32 ? 64 : (NP <= 64 ? 32 : 16));
    __shared__ double s_cq[2 * 256 + 8];
    __shared__ C s_col[2][NP];
    __shared__ C s_row[2][NP];
    __shared__ double2 s_d[2];
    __shared__ int s_perm[NP];
    __shared__ int s_iperm[NP];
    __shared__ int s_sing;
    __shared__ C s_stage[STAGE_ROWS * NP];

    const int b = blockIdx.x, tid = threadIdx.x;
    const int r = tid & 15, c = tid >> 4, col0 = c * TC;
    C *G = gss + (size_t)b * NP * NP;

    if (cqv.cq) rg_load_cq(s_cq, cqv, b, tid, NT);
    if (tid == 0) s_sing = 0;
    __syncthreads();

    // M = (diag(a) - V) - B Gss B in FP64, rounded to T; identity on the padding.
    C a[TR][TC];
#pragma unroll
    for (int i = 0; i < TR; i++) {
        const int gi = r + 16 * i;
#pragma unroll
        for (int j = 0; j < TC; j++) {
            const int gj = col0 + j;
            double2 m;
            if (gi < n && gj < n) {
                m = gi == gj ? adiag[gi] : cmk(0.0, 0.0);
                if (cqv.cq) m = csub(m, slice_v3(s_cq, boundtype, gi, gj));
                if (!first) m = csub(m, cmul(cmul(bvec[gi], ldc2(G + (size_t)gi * NP + gj)), bvec[gj]));
            } else {
                m = cmk(gi == gj ? 1.0 : 0.0, 0.0);
            }
            a[i][j] = c_from<C>(m);
        }
    }

    unsigned pivoted = 0u;
    if (c == 0) pivot_column<0, TR, C>(a, r, pivoted, 0, s_col[0], s_perm, s_d[0], s_sing);
    __syncthreads();
    // Step k = kb + JJ; JJ and the look-ahead column are compile-time so the tile stays in registers.
    auto step = [&](int kb, auto jjc) __attribute__((always_inline)) {
        constexpr int jj = decltype(jjc)::value, jn = (jj + 1) % TC;
        const int k = kb + jj;
        if (k >= n) return;
        const int buf = k & 1;
        const int p = s_perm[k];
        // scaled pivot row from its 32 owners
        if (r == (p & 15)) {
            store_row<0, TR, C>(a, p >> 4, &s_row[buf][col0]);
            pivoted |= 1u << (p >> 4);
        }
        __syncthreads();
        const C dk = c_from<C>(s_d[buf]);
        C rv[TC], m[TR];
#pragma unroll
        for (int j = 0; j < TC; j++) rv[j] = c_mul(s_row[buf][col0 + j], dk);
#pragma unroll
        for (int i = 0; i < TR; i++) m[i] = s_col[buf][r + 16 * i];
        const int kn = k + 1;
        if (kn < n && c == kn / TC) {
            // owners of column k+1: that column first, then the pivot of step k+1
#pragma unroll
            for (int i = 0; i < TR; i++) a[i][jn] = c_fnms(a[i][jn], m[i], rv[jn]);
            pivot_column<jn, TR, C>(a, r, pivoted, kn, s_col[buf ^ 1], s_perm, s_d[buf ^ 1], s_sing);
            // multipliers and row again from LDS: cheaper than keeping them live through the search
#pragma unroll
            for (int j = 0; j < TC; j++) rv[j] = c_mul(s_row[buf][col0 + j], dk);
#pragma unroll
            for (int i = 0; i < TR; i++) m[i] = s_col[buf][r + 16 * i];
#pragma unroll
            for (int i = 0; i < TR; i++)
#pragma unroll
                for (int j = 0; j < TC; j++)
                    if (j != jn) a[i][j] = c_fnms(a[i][j], m[i], rv[j]);
        } else {
#pragma unroll
            for (int i = 0; i < TR; i++)
#pragma unroll
                for (int j = 0; j < TC; j++) a[i][j] = c_fnms(a[i][j], m[i], rv[j]);
        }
        __syncthreads();
    };
    for (int kb = 0; kb < n; kb += TC) {
        step(kb, std::integral_constant<int, 0>{});
        step(kb, std::integral_constant<int, 1>{});
        step(kb, std::integral_constant<int, 2>{});
        step(kb, std::integral_constant<int, 3>{});
    }
    for (int k = n + tid; k < NP; k += NT) s_perm[k] = k;
    __syncthreads();
    for (int k = tid; k < NP; k += NT) s_iperm[s_perm[k]] = k;
    if (tid == 0 && s_sing) status[b] = RG_ESINGULAR;

    // Write inv(M): physical row r' -> iperm[r'], column j <- physical column iperm[j].
    for (int ch = 0; ch < NP; ch += STAGE_ROWS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TR; i++) {
            const int gi = r + 16 * i;
            if (gi >= ch && gi < ch + STAGE_ROWS)
#pragma unroll
                for (int j = 0; j < TC; j++) s_stage[(gi - ch) * NP + col0 + j] = a[i][j];
        }
        __syncthreads();
        for (int e = tid; e < STAGE_ROWS * NP; e += NT) {
            const int rr = e / NP, j = e % NP;
            if (ch + rr >= NP) break;
            G[(size_t)(s_iperm[ch + rr]) * NP + j] = s_stage[rr * NP + s_iperm[j]];
        }
    }
}

}  // namespace

template <class T>
static void launch_gj(rg_plan *P, int nb, typename CT<T>::C *g, const rg_cqv cq, const double2 *ad, int first,
                      const int32_t *only)
{
    switch (P->np / 16) {
#define RG_GJ(W)                                                                                                \
    case W:                                                                                                     \
        hipLaunchKernelGGL((k_gj_lookahead<W, T>), dim3(nb), dim3(64 * W), 0, P->stream, g, cq,                   \
                           P->p.boundtype, ad, P->d_b, P->nmu, first, P->d_status, only, P->cstride, P->d_nx_r, P->cur_ix); \
        break;
        RG_GJ(1)
        RG_GJ(2)
        RG_GJ(3)
        RG_GJ(4)
        RG_GJ(5)
        RG_GJ(6)
        RG_GJ(7)
        RG_GJ(8)
        RG_GJ(9)
        RG_GJ(10)
        RG_GJ(11)
        RG_GJ(12)
        RG_GJ(13)
        RG_GJ(14)
        RG_GJ(15)
        RG_GJ(16)
#undef RG_GJ
        default: break;
    }
}

void rg_launch_inverse_gj(rg_plan *P, int nb, int slot, int first, bool with_disorder, int precision,
                          const int32_t *only)
{
    hipEvent_t stop = nullptr;
    rg_timer_begin(P, only ? RGK_REDO : RGK_INVERSE, &stop);
    const rg_cqv cq = rg_cq_view(P, nb, with_disorder);
    const double2 *ad = P->d_adiag + (size_t)slot * P->np;
    if (precision == RG_FP64) launch_gj<double>(P, nb, P->d_gss, cq, ad, first, only);
    else launch_gj<float>(P, nb, P->d_gss32, cq, ad, first, only);
    rg_timer_end(P, only ? RGK_REDO : RGK_INVERSE, stop);
}
