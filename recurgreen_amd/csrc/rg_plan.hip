// rg_plan.hip -- C ABI of the gfx950 RGF library: plans, batched sweeps, histogram, timing.
//
// A plan is the device-side counterpart of waveguide_t (src/waveguide.f90:13-73): the mode
// constants of init (src/waveguide.f90:81-255) are computed once on the host in FP64, uploaded,
// and every batch buffer is allocated at creation, so the sweep entry points never allocate.
#include <algorithm>
#include <array>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <new>
#include <thread>

#include "rg_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
using zc = std::complex<double>;

inline double2 to2(zc z) { return make_double2(z.real(), z.imag()); }

#define HIPCHK(expr)                                  \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return RG_EDEVICE;      \
    } while (0)

template <class T>
rg_status dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) return RG_OK;
    if (hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count) != hipSuccess) {
        *p = nullptr;
        return RG_ENOMEM;
    }
    return RG_OK;
}

void free_graphs(rg_plan *P);

void free_plan(rg_plan *P)
{
    if (!P) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(P->device);
    void *bufs[] = {P->d_b, P->d_adiag, P->d_sqrtmu, P->d_tab, P->d_gss, P->d_gs1[0], P->d_gs1[1],
                    P->d_tmat, P->d_cq, P->d_teig, P->d_upot, P->d_status, P->d_hist, P->d_bins,
                    P->d_gss32, P->d_gs1_32[0], P->d_gs1_32[1], P->d_g11, P->d_g1s[0], P->d_g1s[1],
                    P->d_sqrtremu, P->d_redo, P->d_us, P->d_ck, P->d_seg, P->d_adj, P->d_loss, P->d_grad, P->d_bconj,
                    P->d_schur, P->d_schur_zero, P->d_schur_i, P->d_tp_g, P->d_tp_z, P->d_tp_x, P->d_tp_v,
                    P->d_tp_w, P->d_tp_y[0], P->d_tp_y[1], P->d_tp_t, P->d_tp_i, P->d_tp_psi, P->d_adj_work,
                    P->d_ref_g, P->d_ref_x, P->d_ref_r, P->d_ref_w, P->d_ref_v, P->d_ref_cq};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    free_graphs(P);
    if (P->d_seed) (void)hipFree(P->d_seed);
    for (auto &t : P->timers) {
        for (auto &e : t.pending) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        for (auto &e : t.pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    }
    if (P->stream) (void)hipStreamDestroy(P->stream);
    for (hipEvent_t e : P->lane_ev)
        if (e) (void)hipEventDestroy(e);
    for (rg_plan *L : P->lane) {
        if (!L) continue;  // a lane shares the plan's buffers: only its stream and events are its own
        for (auto &t : L->timers) {
            for (auto &e : t.pending) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
            for (auto &e : t.pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        }
        if (L->stream) (void)hipStreamDestroy(L->stream);
        delete L;
    }
    (void)hipSetDevice(cur);
    delete P;
}

// Mode set-up, src/waveguide.f90:81-255 (validation :91-140, Dirichlet :191-229, periodic
// :151-189, mu/DOS/B :240-248, slice diagonals :358-363, disorder strength :370).
rg_status init_modes(rg_plan *P)
{
    const rg_params &p = P->p;
    if (!(p.wol >= 2.220446049250313e-16) || !(p.aratio >= 2.220446049250313e-16)) return RG_EINVAL;
    if (p.exfac < 1.0 || p.exfac > 1e4) return RG_EINVAL;
    if (p.nxfree < 0 || p.nxdiso < 1 || p.nxdiso + 2 * p.nxfree < 2) return RG_EINVAL;
    if (p.dscat < 0.0 || p.dabso < 0.0) return RG_EINVAL;
    if (p.napera <= 0.0 || p.napera > 1.0 || p.naperb <= 0.0 || p.naperb > 1.0) return RG_EINVAL;
    P->nxdiso = p.nxdiso;
    P->nxfree = p.nxfree;
    P->nx = p.nxdiso + 2 * p.nxfree;
    P->kl = 2.0 * kPi * p.wol * p.aratio;
    P->kdx = P->kl / p.nxdiso;
    std::vector<double> kp;
    if (p.boundtype == RG_DIRICHLET) {
        if (std::fmod(p.wol, 0.5) == 0.0) return RG_EINVAL;
        P->nmu = (int)std::floor(2.0 * p.exfac * p.wol);
        P->imin = 1;
        P->imax = (int)std::floor(2.0 * p.naperb * p.wol);
        P->jmin = 1;
        P->jmax = (int)std::floor(2.0 * p.napera * p.wol);
        P->neigv = (int)std::floor(2.0 * std::min(p.napera, p.naperb) * p.wol);
        P->nprop = (int)std::floor(2.0 * p.wol);
        for (int m = 1; m <= P->nmu; m++) kp.push_back(0.5 * m / p.wol);
    } else if (p.boundtype == RG_PERIODIC) {
        if (std::fmod(p.wol, 1.0) == 0.0) return RG_EINVAL;
        P->nmu = 2 * (int)std::floor(p.exfac * p.wol) + 1;
        P->imin = P->nmu / 2 + 1 - (int)std::floor(p.naperb * p.wol);
        P->imax = P->nmu / 2 + 1 + (int)std::floor(p.naperb * p.wol);
        P->jmin = P->nmu / 2 + 1 - (int)std::floor(p.napera * p.wol);
        P->jmax = P->nmu / 2 + 1 + (int)std::floor(p.napera * p.wol);
        P->neigv = 2 * (int)std::floor(std::min(p.napera, p.naperb) * p.wol) + 1;
        P->nprop = 2 * (int)std::floor(p.wol) + 1;
        for (int m = 1; m <= P->nmu; m++) kp.push_back((double)(m - P->nmu / 2 - 1) / p.wol);
    } else {
        return RG_EINVAL;
    }
    P->ny = p.ny > 0 ? p.ny : P->nmu;
    if (P->nmu < 1 || P->ny < P->nmu) return RG_EINVAL;
    if (P->neigv < 1) return RG_EINVAL;
    const int n = P->nmu;
    P->np = ((n + 15) / 16) * 16;
    double dsum = 0.0;
    P->h_mu.assign(P->np, make_double2(0, 0));
    P->h_b.assign(P->np, make_double2(0, 0));
    P->h_abound.assign(P->np, make_double2(0, 0));
    P->h_abulk.assign(P->np, make_double2(0, 0));
    P->h_sqrtmu.assign(P->np, make_double2(0, 0));
    for (int m = 0; m < n; m++) {
        const double kp2 = kp[m] * kp[m];
        const zc mu = std::sqrt(zc(1.0 - kp2, p.dabso / P->kl));
        dsum += (1.0 / std::sqrt(zc(1.0 - kp2, 0.0))).real();
        const zc th = mu * P->kdx;
        P->h_mu[m] = to2(mu);
        P->h_b[m] = to2(mu / std::sin(th));
        P->h_abound[m] = to2(zc(0.0, 1.0) * mu - mu / std::tan(th));
        P->h_abulk[m] = to2(-2.0 * mu / std::tan(th));
        P->h_sqrtmu[m] = to2(std::sqrt(mu));
    }
    P->dos = p.dos_model == 1 ? 1.0 / (4.0 * kPi) : dsum / (p.wol * (2.0 * kPi) * (2.0 * kPi));
    P->stdu = std::sqrt(p.dscat * P->ny / (2.0 * kPi * kPi * P->dos * p.wol * p.nxdiso));
    P->ntab = p.boundtype == RG_DIRICHLET ? 2 * n + 1 : 2 * n;
    P->ntp = (P->ntab + 15) / 16 * 16;
    P->nchunk = (P->ny + 15) / 16;
    P->cq_parts = rg_project_parts(P->ny);
    return RG_OK;
}

rg_status upload_constants(rg_plan *P)
{
    const int np = P->np, ny = P->ny, nt = P->ntab;
    std::vector<double2> ad(3 * (size_t)np, make_double2(0, 0));
    for (int m = 0; m < P->nmu; m++) {
        ad[m] = P->h_abound[m];
        ad[np + m] = P->h_abulk[m];
        // last slice: boundary diagonal; when nx == 2 the reference adds the bulk diagonal to
        // slice 2 after copying the boundary (src/waveguide.f90:359-360).
        ad[2 * np + m] = P->nx == 2 ? cadd(P->h_abound[m], P->h_abulk[m]) : P->h_abound[m];
    }
    // projection table, rows = transverse cells (zero rows up to whole 16-cell chunks), columns
    // padded to whole 16-column tiles
    const int ntp = P->ntp, rows = 16 * P->nchunk;
    std::vector<double> tab((size_t)rows * ntp, 0.0);
    for (int i = 0; i < ny; i++) {
        const double y = (i + 1 - 0.5) / ny;
        if (P->p.boundtype == RG_DIRICHLET) {
            for (int q = 0; q < nt; q++) tab[(size_t)i * ntp + q] = std::cos(q * kPi * y);
        } else {
            for (int q = 0; q < nt / 2; q++) {
                tab[(size_t)i * ntp + 2 * q] = std::cos(2.0 * kPi * q * y);
                tab[(size_t)i * ntp + 2 * q + 1] = std::sin(2.0 * kPi * q * y);
            }
        }
    }
    rg_status st;
    if ((st = dalloc(&P->d_b, np)) || (st = dalloc(&P->d_adiag, 3 * (size_t)np)) || (st = dalloc(&P->d_sqrtmu, np)) ||
        (st = dalloc(&P->d_tab, tab.size())))
        return st;
    HIPCHK(hipMemcpy(P->d_b, P->h_b.data(), sizeof(double2) * np, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P->d_adiag, ad.data(), sizeof(double2) * ad.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P->d_sqrtmu, P->h_sqrtmu.data(), sizeof(double2) * np, hipMemcpyHostToDevice));
    std::vector<double2> bc(np, make_double2(0, 0));
    for (int m = 0; m < P->nmu; m++) bc[m] = cconj(P->h_b[m]);
    if ((st = dalloc(&P->d_bconj, np))) return st;
    HIPCHK(hipMemcpy(P->d_bconj, bc.data(), sizeof(double2) * np, hipMemcpyHostToDevice));
    std::vector<double> sre(np, 0.0);
    for (int m = 0; m < P->nmu; m++) sre[m] = std::sqrt(P->h_mu[m].x);  // src/waveguide.f90:510
    if ((st = dalloc(&P->d_sqrtremu, np))) return st;
    HIPCHK(hipMemcpy(P->d_sqrtremu, sre.data(), sizeof(double) * np, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P->d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    return RG_OK;
}

rg_status alloc_batch(rg_plan *P)
{
    const size_t B = (size_t)P->batch_max, mat = (size_t)P->np * P->np;
    rg_status st;
    if ((st = dalloc(&P->d_gss, B * mat)) || (st = dalloc(&P->d_gs1[0], B * mat)) || (st = dalloc(&P->d_gs1[1], B * mat)) ||
        (st = dalloc(&P->d_tmat, B * mat)) || (st = dalloc(&P->d_cq, B * P->ntab * kMaxCqParts)) ||
        (st = dalloc(&P->d_teig, B * P->neigv)) || (st = dalloc(&P->d_status, B)) || (st = dalloc(&P->d_redo, B)) || (st = dalloc(&P->d_us, B * P->ny)) ||
        (st = dalloc(&P->d_gss32, B * mat)) || (st = dalloc(&P->d_gs1_32[0], B * mat)) ||
        (st = dalloc(&P->d_gs1_32[1], B * mat)))
        return st;
    const size_t per = (size_t)P->nxdiso * P->ny;
    const size_t cap_bytes = (size_t)256 << 20;
    P->upot_cap = (int)std::max<size_t>(1, std::min<size_t>(B, cap_bytes / (per * sizeof(double))));
    if ((st = dalloc(&P->d_upot, (size_t)P->upot_cap * per))) return st;
    P->refl_cap = (int)std::min<size_t>(B, 256);
    if ((st = dalloc(&P->d_g11, (size_t)P->refl_cap * mat)) || (st = dalloc(&P->d_g1s[0], (size_t)P->refl_cap * mat)) ||
        (st = dalloc(&P->d_g1s[1], (size_t)P->refl_cap * mat)))
        return st;
    if (P->np > 128) {  // block work area of the Schur-complement inversion
        if ((st = dalloc(&P->d_schur, B * rg_schur_elems(P->np))) || (st = dalloc(&P->d_schur_zero, 3 * 128)) ||
            (st = dalloc(&P->d_schur_i, 128)))
            return st;
        std::vector<double2> zero(3 * 128, make_double2(0, 0)), iv(128, make_double2(0, 1));
        HIPCHK(hipMemcpy(P->d_schur_zero, zero.data(), sizeof(double2) * zero.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(P->d_schur_i, iv.data(), sizeof(double2) * iv.size(), hipMemcpyHostToDevice));
    }
    P->hist_cap = 4096;
    if ((st = dalloc(&P->d_hist, (size_t)P->hist_cap + 2)) || (st = dalloc(&P->d_bins, (size_t)P->hist_cap + 1))) return st;
    return RG_OK;
}

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) { (void)hipGetDevice(&prev); (void)hipSetDevice(dev); }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

void sweep_step(rg_plan *P, int nb, int ix, int slot, int ixd, int64_t seed0, const double *d_upot, int precision,
                int *cur, bool flip);

// One slice of the sweep over nb realizations (src/waveguide.f90:424-440): projection of the
// slice's disorder, inversion of the slice matrix into Gss, and Gs1 <- -Gss B Gs1 (a copy of Gss
// on the first slice).  *cur tracks which Gs1 buffer holds the state.  FP64 keeps complex128
// state; the mixed modes keep complex64 state with FP64 projection, pivot checks and eigenvalues.
void sweep_slice(rg_plan *P, int nb, int ix, int64_t seed0, const double *d_upot, int precision, int *cur)
{
    const int id = ix - P->nxfree;  // 1-based index in the disordered region
    const bool dis = id >= 1 && id <= P->nxdiso;
    sweep_step(P, nb, ix, ix == 1 ? 0 : (ix == P->nx ? 2 : 1), dis ? id - 1 : -1, seed0, d_upot, precision, cur, true);
}

// The slice-inversion kernel of a sweep in `precision`: the plan's choice, where it applies.
int inverse_kernel(const rg_plan *P, int precision)
{
    int inv = precision == RG_FP64 ? P->inv_fp64 : P->inv_mixed;
    if (inv == RG_INV_SYM && P->p.boundtype != RG_DIRICHLET) inv = RG_INV_NOPIV;  // periodic: V is Hermitian
    if ((inv == RG_INV_NOPIV || inv == RG_INV_SYM) && P->np < 32) inv = RG_INV_LOOKAHEAD;
    return inv;
}

// Orders above 128 invert through the Schur complement, except complex64 sweeps with symmetric slice
// matrices, which the symmetric kernel inverts directly (one wave per 16-column block) with the pivot
// checks over the whole matrix: the Schur complement keeps the pivoting inside its two blocks, and in
// FP32 the complement D - C A^-1 B loses what zgetrf's pivoting across the blocks would keep (a
// potential spike of 1e5 standard deviations: T off by 4e-2 at N = 256).
bool schur_inversion(const rg_plan *P, int precision)
{
    if (P->np <= 128) return false;
    return !(precision != RG_FP64 && inverse_kernel(P, precision) == RG_INV_SYM);
}

// One slice step with the slice's diagonal (slot 0 first, 1 bulk, 2 last) and disordered-slice index
// ixd (-1: no disorder) given; flip = false leaves *cur for the caller to flip (several ranges of
// one batch stepping together).
void sweep_step(rg_plan *P, int nb, int ix, int slot, int ixd, int64_t seed0, const double *d_upot, int precision,
                int *cur, bool flip)
{
    const size_t mat = (size_t)nb * P->np * P->np;
    const bool dis = ixd >= 0;
    if (dis) rg_launch_project(P, nb, ixd, seed0, d_upot, nb);
    int inv = inverse_kernel(P, precision);
    if (schur_inversion(P, precision)) rg_launch_inverse_schur(P, nb, slot, ix == 1, dis, precision);
    else if (inv == RG_INV_REGISTER) rg_launch_inverse_f64(P, nb, slot, ix == 1, dis);
    else if (inv == RG_INV_LOOKAHEAD) rg_launch_inverse_gj(P, nb, slot, ix == 1, dis, precision);
    else if (inv == RG_INV_NOPIV || inv == RG_INV_SYM) {
        if (precision == RG_FP64) {
            if (inv == RG_INV_SYM) rg_launch_inverse_sym64(P, nb, slot, ix == 1, dis, P->d_redo);
            else rg_launch_inverse_nopiv64(P, nb, slot, ix == 1, dis, P->d_redo);
        } else {
            if (inv == RG_INV_SYM) rg_launch_inverse_sym32(P, nb, slot, ix == 1, dis, P->d_redo);
            else rg_launch_inverse_nopiv(P, nb, slot, ix == 1, dis, P->d_redo);
        }
        rg_launch_inverse_gj(P, nb, slot, ix == 1, dis, precision, P->d_redo);
    } else rg_launch_inverse_blocked(P, nb, slot, ix == 1, dis, precision);
    hipEvent_t stop = nullptr;
    if (precision == RG_FP64) {
        if (ix == 1) {  // a copy kernel (complex128 as pairs of float2), so captured sweeps hold kernels only
            rg_timer_begin(P, RGK_COPY, &stop);
            rg_launch_copy_f2(P, reinterpret_cast<const float2 *>(P->d_gss), reinterpret_cast<float2 *>(P->d_gs1[0]),
                              2 * mat);
            rg_timer_end(P, RGK_COPY, stop);
        } else {
            rg_launch_gemm_f64(P, nb, P->d_gss, P->d_gs1[*cur], P->d_gs1[*cur ^ 1]);
            if (flip) *cur ^= 1;
        }
    } else {
        if (ix == 1) {
            rg_timer_begin(P, RGK_COPY, &stop);
            rg_launch_copy_f2(P, P->d_gss32, P->d_gs1_32[0], mat);
            rg_timer_end(P, RGK_COPY, stop);
        } else {
            rg_launch_gemm_sp(P, nb, P->d_gss32, P->d_gs1_32[*cur], P->d_gs1_32[*cur ^ 1], precision);
            if (flip) *cur ^= 1;
        }
    }
}

// After the last slice: the Fisher-Lee relation and the transmission eigenvalues (:443-449) into
// teig, or the t matrices into tmat_out.
void sweep_finish(rg_plan *P, int nb, int precision, int cur, double *teig, double2 *tmat_out)
{
    if (tmat_out) {
        if (precision == RG_FP64) rg_launch_fisher_lee_svd_f64(P, nb, P->d_gs1[cur], teig, P->d_status, tmat_out);
        else rg_launch_fisher_lee_svd_f32(P, nb, P->d_gs1_32[cur], teig, P->d_status, tmat_out);
    } else if (precision == RG_FP64) {
        rg_launch_teig_f64(P, nb, P->d_gs1[cur], teig, P->d_status);
    } else {
        rg_launch_teig_f32(P, nb, P->d_gs1_32[cur], teig, P->d_status);
    }
}

// Lane k of a split batch: the realizations [off, off + nb) of the plan's buffers, launched on the
// lane's own stream.  Every pointer the sweep touches is re-based on each call.
static void lane_view(const rg_plan *P, rg_plan *L, int off, int nb)
{
    const size_t mat = (size_t)P->np * P->np;
    L->d_gss = P->d_gss + off * mat;
    L->d_gss32 = P->d_gss32 + off * mat;
    for (int i = 0; i < 2; i++) {
        L->d_gs1[i] = P->d_gs1[i] + off * mat;
        L->d_gs1_32[i] = P->d_gs1_32[i] + off * mat;
    }
    // partial sums [part][nb][ntab]: lane 1's block follows lane 0's (cq_parts * off * ntab)
    L->d_cq = P->d_cq + (size_t)P->cq_parts * off * P->ntab;
    L->d_us = P->d_us + (size_t)off * P->ny;
    L->d_status = P->d_status + off;
    L->d_redo = P->d_redo + off;
    if (P->cstride) {  // constants of each realization
        L->d_adiag = P->d_adiag + (size_t)off * 3 * P->cstride;
        L->d_b = P->d_b + (size_t)off * P->cstride;
        L->d_sqrtmu = P->d_sqrtmu + (size_t)off * P->cstride;
    }
    L->cstride = P->cstride;
    L->d_seeds_r = P->d_seeds_r ? P->d_seeds_r + off : nullptr;
    L->d_nx_r = P->d_nx_r ? P->d_nx_r + off : nullptr;
    L->schur_off = P->schur_off + off;
    L->cur_ix = P->cur_ix;
    L->d_stdu_r = P->d_stdu_r ? P->d_stdu_r + off : nullptr;
    L->inv_fp64 = P->inv_fp64;
    L->inv_mixed = P->inv_mixed;
    L->gemm_wg = P->gemm_wg;
    L->dyson64 = P->dyson64;
    L->project_wide_min = P->project_wide_min;
    L->eigen = P->eigen;
    L->profiling = P->profiling;
    L->profile_mask = P->profile_mask;
    (void)nb;
}

namespace {
__global__ void k_set_seed(int64_t *__restrict__ dst, int64_t v) { *dst = v; }
}  // namespace

// Options that change the kernels a sweep launches (part of a captured sweep's key).
static int graph_opts(const rg_plan *P)
{
    return (P->inv_fp64 | P->inv_mixed << 4 | P->cq_parts << 9 | P->eigen << 20 | P->dyson64 << 14 |
            P->gemm_wg << 16) ^ (int)((uint32_t)P->project_wide_min * 0x9E3779B1u & 0x7F000000u);
}

void free_graphs(rg_plan *P)
{
    for (auto &e : P->graphs) {
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
        if (e.graph) (void)hipGraphDestroy(e.graph);
    }
    P->graphs.clear();
}

// The sweep as one executable graph: captured on first use for (nb, precision, potentials,
// output, options), then replayed.  The seed is written to the device word d_seed by a kernel
// launched before the replay, so one graph serves every seed range.  Returns false (nothing
// launched) if the capture or instantiation fails; the caller then launches the sweep directly.
bool run_sweep_graph(rg_plan *P, int nb, int64_t seed0, const double *d_upot, int precision, double *teig)
{
    if (!P->d_seed && hipMalloc(&P->d_seed, sizeof(int64_t)) != hipSuccess) {
        P->d_seed = nullptr;
        return false;
    }
    const int opts = graph_opts(P);
    rg_plan::graph_entry *hit = nullptr;
    for (auto &e : P->graphs)
        if (e.nb == nb && e.precision == precision && e.opts == opts && e.upot == d_upot && e.teig == teig) hit = &e;
    if (!hit) {
        if (P->graphs.size() >= 8) {  // keep a few shapes: drop the oldest
            (void)hipGraphExecDestroy(P->graphs.front().exec);
            (void)hipGraphDestroy(P->graphs.front().graph);
            P->graphs.erase(P->graphs.begin());
        }
        rg_plan::graph_entry e{nb, precision, opts, d_upot, teig, nullptr, nullptr};
        if (hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
        P->seed_dev = P->d_seed;
        int cur = 0;
        for (int ix = 1; ix <= P->nx; ix++) sweep_slice(P, nb, ix, 0, d_upot, precision, &cur);
        sweep_finish(P, nb, precision, cur, teig, nullptr);
        P->seed_dev = nullptr;
        const hipError_t ce = hipStreamEndCapture(P->stream, &e.graph);
        if (ce != hipSuccess || !e.graph) return false;
        if (hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(e.graph);
            return false;
        }
        P->graphs.push_back(e);
        hit = &P->graphs.back();
    }
    hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(1), 0, P->stream, P->d_seed, seed0);
    return hipGraphLaunch(hit->exec, P->stream) == hipSuccess;
}

// The sweep over nb realizations into teig (and tmat_out when given).  With two lanes the batch is
// split in halves on two streams, slice launches interleaved and the second lane started one slice
// inversion behind the first, so that one half's latency-bound inversion shares the compute units
// with the other half's Dyson products.
// A complex64 sweep with FP64 refinement: the sweep keeps every slice's inverse and coefficients, the
// refinement passes (rg_refine.hip) give G(L;0) to FP64 accuracy, and the FP64 Fisher-Lee relation and
// eigenvalues follow.  nb <= ref_cap.
static void run_sweep_refined(rg_plan *P, int nb, int64_t seed0, const double *d_upot, int precision, double *teig,
                       double2 *tmat_out)
{
    int cur = 0;
    for (int ix = 1; ix <= P->nx; ix++) {
        sweep_slice(P, nb, ix, seed0, d_upot, precision, &cur);
        const int id = ix - P->nxfree;
        rg_refine_store(P, nb, ix, id >= 1 && id <= P->nxdiso, P->d_gs1_32[cur]);
    }
    const double2 *gl0 = rg_refine_solve(P, nb);
    if (tmat_out) rg_launch_fisher_lee_svd_f64(P, nb, gl0, teig, P->d_status, tmat_out);
    else rg_launch_teig_f64(P, nb, gl0, teig, P->d_status);
}

static bool refined(const rg_plan *P, int precision) { return P->refine > 0 && precision != RG_FP64; }

// Realizations per pass: `base`, or fewer where the refinement's per-slice store sets the limit.
static rg_status pass_cap(rg_plan *P, int precision, int base, int *cap)
{
    *cap = base;
    if (!refined(P, precision)) return RG_OK;
    rg_status st = rg_refine_reserve(P);
    if (st) return st;
    *cap = std::min(base, P->ref_cap);
    return RG_OK;
}

void run_sweep(rg_plan *P, int nb, int64_t seed0, const double *d_upot, int precision, double *teig,
               double2 *tmat_out)
{
    if (refined(P, precision)) {
        run_sweep_refined(P, nb, seed0, d_upot, precision, teig, tmat_out);
        return;
    }
    if (P->use_graph && !P->profiling && !tmat_out && (P->lanes < 2 || nb < 2 * kLaneMin) &&
        run_sweep_graph(P, nb, seed0, d_upot, precision, teig))
        return;
    if (P->lanes < 2 || tmat_out || nb < 2 * kLaneMin || P->np > 128) {
        int cur = 0;
        for (int ix = 1; ix <= P->nx; ix++) sweep_slice(P, nb, ix, seed0, d_upot, precision, &cur);
        sweep_finish(P, nb, precision, cur, teig, tmat_out);
        return;
    }
    const int h = (nb + 1) / 2, n[2] = {h, nb - h}, off[2] = {0, h};
    const size_t per = (size_t)P->nxdiso * P->ny;
    int cur[2] = {0, 0};
    (void)hipEventRecord(P->lane_ev[0], P->stream);
    for (int k = 0; k < 2; k++) {
        lane_view(P, P->lane[k], off[k], n[k]);
        (void)hipStreamWaitEvent(P->lane[k]->stream, P->lane_ev[0], 0);
    }
    for (int ix = 1; ix <= P->nx; ix++)
        for (int k = 0; k < 2; k++) {
            rg_plan *L = P->lane[k];
            sweep_slice(L, n[k], ix, seed0 + off[k], d_upot ? d_upot + off[k] * per : nullptr, precision, &cur[k]);
            if (ix == 1 && k == 0) {
                (void)hipEventRecord(P->lane_ev[1], L->stream);
                (void)hipStreamWaitEvent(P->lane[1]->stream, P->lane_ev[1], 0);
            }
        }
    for (int k = 0; k < 2; k++) {
        rg_plan *L = P->lane[k];
        sweep_finish(L, n[k], precision, cur[k], teig + (size_t)off[k] * P->neigv, nullptr);
        (void)hipEventRecord(P->lane_ev[1 + k], L->stream);
        (void)hipStreamWaitEvent(P->stream, P->lane_ev[1 + k], 0);
    }
}

// complex128 slice inversion of the reflection and adjoint sweeps: the plan's kernel choice, or the
// Schur complement above np = 128.
void invert_f64(rg_plan *P, int nb, int slot, int first, bool dis)
{
    if (P->np > 128) rg_launch_inverse_schur(P, nb, slot, first, dis, RG_FP64);
    else {
        const int inv = inverse_kernel(P, RG_FP64);
        if (inv == RG_INV_REGISTER) rg_launch_inverse_f64(P, nb, slot, first, dis);
        else if (inv == RG_INV_LOOKAHEAD) rg_launch_inverse_gj(P, nb, slot, first, dis, RG_FP64);
        else if (inv == RG_INV_NOPIV || inv == RG_INV_SYM) {
            if (inv == RG_INV_SYM) rg_launch_inverse_sym64(P, nb, slot, first, dis, P->d_redo);
            else rg_launch_inverse_nopiv64(P, nb, slot, first, dis, P->d_redo);
            rg_launch_inverse_gj(P, nb, slot, first, dis, RG_FP64, P->d_redo);
        } else rg_launch_inverse_blocked(P, nb, slot, first, dis, RG_FP64);
    }
}

// The reflection/transmission sweep of compute_refl_trans_v1 (src/waveguide.f90:469-519), FP64:
// besides Gss and Gs1 it carries G11 and G1s, updated after Gs1 in the reference's order
// (:498-505): Gs1 <- -Gss B Gs1, G11 <- G11 - G1s B Gs1 (new Gs1), G1s <- -G1s B Gss.
// Writes r (column-major nmu x nmu per realization) to r_out and t to the Gs1 buffer the sweep
// finished without, which it returns.
double2 *run_sweep_refl(rg_plan *P, int nb, int64_t seed0, const double *d_upot, double2 *r_out)
{
    int cur = 0, c1s = 0;
    const int nx = P->nx;
    const size_t mat = (size_t)nb * P->np * P->np;
    const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
    for (int ix = 1; ix <= nx; ix++) {
        const int id = ix - P->nxfree;
        const bool dis = id >= 1 && id <= P->nxdiso;
        if (dis) rg_launch_project(P, nb, id - 1, seed0, d_upot, nb);
        const int slot = ix == 1 ? 0 : (ix == nx ? 2 : 1);
        invert_f64(P, nb, slot, ix == 1, dis);
        if (ix == 1) {
            (void)hipMemcpyAsync(P->d_gs1[0], P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
            (void)hipMemcpyAsync(P->d_g11, P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
            (void)hipMemcpyAsync(P->d_g1s[0], P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
        } else {
            rg_launch_gemm_f64(P, nb, P->d_gss, P->d_gs1[cur], P->d_gs1[cur ^ 1]);
            cur ^= 1;
            rg_zgemm(P, nb, mone, P->d_g1s[c1s], 0, P->d_gs1[cur], 0, one, P->d_g11, nullptr, P->d_b, nullptr);
            rg_zgemm(P, nb, mone, P->d_g1s[c1s], 0, P->d_gss, 0, zero, P->d_g1s[c1s ^ 1], nullptr, P->d_b, nullptr);
            c1s ^= 1;
        }
    }
    rg_launch_refl_trans_out(P, nb, P->d_g11, P->d_gs1[cur], r_out, P->d_gs1[cur ^ 1]);
    return P->d_gs1[cur ^ 1];
}

// compute_deposition (src/waveguide.f90:710-745), by recursion instead of the reference's
// block-tridiagonal solve.  The slices on either side of fpos fold into the left-connected Green
// function g = (A_x - V_x - B g' B)^-1 (a sweep from slice 1 up to fpos-1) and the right-connected h
// (the same kernels from slice nx down to fpos+1); slice fpos coupled to both gives
// G(fpos; fpos) = (A_f - V_f - B (g + h) B)^-1.  The sweep on the side of ipos also carries the
// propagator of its part from ipos, P_ipos = g_ipos (or h_ipos), P_x = -g_x B P_(x-1) (or
// -h_x B P_(x+1)), as Gs1 of the transmission sweep (src/waveguide.f90:438), and
// G(fpos; ipos) = -G(fpos; fpos) B P at the slice next to fpos (G(fpos; fpos) itself at ipos = fpos).
// The part without ipos is kept in the G11 buffer of the reflection sweep.  Writes
// dep = 2i s G(fpos; ipos) s (column-major nmu x nmu) to dep_out.
void run_sweep_deposition(rg_plan *P, int nb, int64_t seed0, const double *d_upot, int fpos, int ipos,
                          double2 *dep_out)
{
    const int nx = P->nx;
    const size_t mat = (size_t)nb * P->np * P->np;
    auto slot_of = [nx](int ix) { return ix == 1 ? 0 : (ix == nx ? 2 : 1); };
    auto project = [&](int ix) {
        const int id = ix - P->nxfree;
        const bool dis = id >= 1 && id <= P->nxdiso;
        if (dis) rg_launch_project(P, nb, id - 1, seed0, d_upot, nb);
        return dis;
    };
    int cur = 0;
    // one slice of a sweep started at slice `first`: Gss, and the propagator from ipos once reached
    auto step = [&](int ix, int first, bool carry) {
        const bool dis = project(ix);
        invert_f64(P, nb, slot_of(ix), ix == first, dis);
        if (!carry) return;
        if (ix == ipos) {
            (void)hipMemcpyAsync(P->d_gs1[0], P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
            cur = 0;
        } else {
            rg_launch_gemm_f64(P, nb, P->d_gss, P->d_gs1[cur], P->d_gs1[cur ^ 1]);
            cur ^= 1;
        }
    };
    // slice fpos: the inversion subtracts B Gss B with Gss = g + h (h alone at fpos = 1, g alone at
    // fpos = nx; nx >= 2, so fpos always has a neighbour).  The part without ipos goes first and waits
    // in the G11 buffer.
    auto keep = [&]() {
        (void)hipMemcpyAsync(P->d_g11, P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
    };
    if (ipos <= fpos) {
        if (fpos < nx) {
            for (int ix = nx; ix > fpos; ix--) step(ix, nx, false);
            if (fpos > 1) keep();
        }
        if (fpos > 1) {
            for (int ix = 1; ix < fpos; ix++) step(ix, 1, ipos < fpos && ix >= ipos);
            if (fpos < nx) rg_launch_add_c128(P, P->d_gss, P->d_g11, mat);
        }
    } else {
        if (fpos > 1) {
            for (int ix = 1; ix < fpos; ix++) step(ix, 1, false);
            keep();
        }
        for (int ix = nx; ix > fpos; ix--) step(ix, nx, ix <= ipos);
        if (fpos > 1) rg_launch_add_c128(P, P->d_gss, P->d_g11, mat);
    }
    const bool dis = project(fpos);
    invert_f64(P, nb, slot_of(fpos), 0, dis);
    const double2 *gfi = P->d_gss;
    if (ipos != fpos) {
        rg_launch_gemm_f64(P, nb, P->d_gss, P->d_gs1[cur], P->d_gs1[cur ^ 1]);
        gfi = P->d_gs1[cur ^ 1];
    }
    rg_launch_deposition_out(P, nb, gfi, dep_out);
}

// Transmission eigenstates inside the medium (compute_tprofile, src/waveguide.f90:843-907), complex128:
// the transmission sweep keeps Gss_x (g_x) and Gs1_x (z_x) of every slice; t = 2i s z_nx s is
// decomposed as t^H t = U diag(T) U^H; with W = s U the column block Y_x = G(x;1) W follows from the
// back substitution of the block-tridiagonal solve (src/base_utils.f90:330-333),
// Y_nx = z_nx W, Y_x = z_x W - g_x B Y_(x+1), and the intensity of eigenvector v at slice x is
// |Y_x(:, v)|^2.  filtered: t over the apertures (rows imin..imax, columns jmin..jmax) as in
// compute_tprofile_group (:951-954), else the full nmu x nmu matrix (:851).  Writes T (descending,
// nv per realization) to P->d_tp_t and the intensities (nb x nx x nv) to P->d_tp_i.
void run_sweep_tprofile(rg_plan *P, int nb, int64_t seed0, const double *d_upot, bool filtered)
{
    const int nx = P->nx;
    const int i0 = filtered ? P->imin - 1 : 0, m = filtered ? P->imax - P->imin + 1 : P->nmu;
    const int j0 = filtered ? P->jmin - 1 : 0, n = filtered ? P->jmax - P->jmin + 1 : P->nmu;
    const size_t mat = (size_t)nb * P->np * P->np, slab = (size_t)P->tp_cap * P->np * P->np;
    const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
    for (int ix = 1; ix <= nx; ix++) {
        const int id = ix - P->nxfree;
        const bool dis = id >= 1 && id <= P->nxdiso;
        if (dis) rg_launch_project(P, nb, id - 1, seed0, d_upot, nb);
        invert_f64(P, nb, ix == 1 ? 0 : (ix == nx ? 2 : 1), ix == 1, dis);
        double2 *g = P->d_tp_g + (ix - 1) * slab, *z = P->d_tp_z + (ix - 1) * slab;
        (void)hipMemcpyAsync(g, P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
        if (ix == 1) (void)hipMemcpyAsync(z, P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
        else rg_launch_gemm_f64(P, nb, P->d_gss, z - slab, z);
    }
    rg_launch_tprofile_svd(P, nb, P->d_tp_z + (nx - 1) * slab, i0, m, j0, n, P->d_tp_x, P->d_tp_v, P->d_tp_w,
                           P->d_tp_t);
    double2 *ya = P->d_tp_y[0], *yb = P->d_tp_y[1];
    const size_t ostride = (size_t)nx * n;
    rg_zgemm(P, nb, one, P->d_tp_z + (nx - 1) * slab, 0, P->d_tp_w, 0, zero, ya, nullptr, nullptr, nullptr);
    rg_launch_colnorms(P, nb, ya, n, P->d_tp_i + (size_t)(nx - 1) * n, ostride);
    for (int k = nx - 2; k >= 0; k--) {
        rg_zgemm(P, nb, one, P->d_tp_z + k * slab, 0, P->d_tp_w, 0, zero, yb, nullptr, nullptr, nullptr);
        rg_zgemm(P, nb, mone, P->d_tp_g + k * slab, 0, ya, 0, one, yb, nullptr, P->d_b, nullptr);
        rg_launch_colnorms(P, nb, yb, n, P->d_tp_i + (size_t)k * n, ostride);
        std::swap(ya, yb);
    }
}

// compute_wavefunction (src/waveguide.f90:636-697) for one realization: the column imu of G(x; ipos) for
// every slice x by the block-tridiagonal solve's two recursions on the stored slice Green functions
// g_x (src/base_utils.f90:320-333): z_x = 0 (x < ipos), z_ipos = g_ipos e_imu, z_x = -g_x B z_(x-1);
// y_nx = z_nx, y_x = z_x - g_x B y_(x+1); each y_x goes through the transverse modes into psi[x][:]
// (ny complex per slice).  Uses the tprofile workspace (one realization).
void run_sweep_wavefunction(rg_plan *P, int64_t seed0, const double *d_upot, int ipos, int imu, double2 *psi)
{
    const int nx = P->nx, nb = 1;
    const size_t mat = (size_t)P->np * P->np, slab = (size_t)P->tp_cap * mat;
    const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
    for (int ix = 1; ix <= nx; ix++) {
        const int id = ix - P->nxfree;
        const bool dis = id >= 1 && id <= P->nxdiso;
        if (dis) rg_launch_project(P, nb, id - 1, seed0, d_upot, nb);
        invert_f64(P, nb, ix == 1 ? 0 : (ix == nx ? 2 : 1), ix == 1, dis);
        (void)hipMemcpyAsync(P->d_tp_g + (ix - 1) * slab, P->d_gss, sizeof(double2) * mat, hipMemcpyDeviceToDevice,
                             P->stream);
    }
    double2 *g = P->d_tp_g, *z = P->d_tp_z;
    rg_launch_unit_column(P, P->d_tp_w, imu - 1);
    rg_zgemm(P, nb, one, g + (ipos - 1) * slab, 0, P->d_tp_w, 0, zero, z + (ipos - 1) * slab, nullptr, nullptr, nullptr);
    for (int ix = ipos + 1; ix <= nx; ix++)
        rg_zgemm(P, nb, mone, g + (ix - 1) * slab, 0, z + (ix - 2) * slab, 0, zero, z + (ix - 1) * slab, nullptr, P->d_b,
                 nullptr);
    double2 *ya = P->d_tp_y[0], *yb = P->d_tp_y[1];
    (void)hipMemcpyAsync(ya, z + (nx - 1) * slab, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
    rg_launch_modal_column(P, ya, psi + (size_t)(nx - 1) * P->ny);
    for (int ix = nx - 1; ix >= 1; ix--) {
        double2 beta = zero;
        if (ix >= ipos) {
            (void)hipMemcpyAsync(yb, z + (ix - 1) * slab, sizeof(double2) * mat, hipMemcpyDeviceToDevice, P->stream);
            beta = one;
        }
        rg_zgemm(P, nb, mone, g + (ix - 1) * slab, 0, ya, 0, beta, yb, nullptr, P->d_b, nullptr);
        rg_launch_modal_column(P, yb, psi + (size_t)(ix - 1) * P->ny);
        std::swap(ya, yb);
    }
}

// One complex128 slice step: Gss_x in place in P->d_gss, and s1_out = -Gss_x B s1_in (a copy of
// Gss_1 at ix == 1).  d_upot: explicit potentials of the pass (nullptr: no disorder).
void fwd_step_f64(rg_plan *P, int nb, int ix, const double *d_upot, const double2 *s1_in, double2 *s1_out)
{
    const int id = ix - P->nxfree;
    const bool dis = d_upot && id >= 1 && id <= P->nxdiso;
    if (dis) rg_launch_project(P, nb, id - 1, 0, d_upot, nb);
    const int slot = ix == 1 ? 0 : (ix == P->nx ? 2 : 1);
    invert_f64(P, nb, slot, ix == 1, dis);
    if (ix == 1)
        (void)hipMemcpyAsync(s1_out, P->d_gss, sizeof(double2) * (size_t)nb * P->np * P->np, hipMemcpyDeviceToDevice,
                             P->stream);
    else
        rg_launch_gemm_f64(P, nb, P->d_gss, s1_in, s1_out);
}

// The same step on the complex64 state (the kernels of the mixed sweep).
void fwd_step_f32(rg_plan *P, int nb, int ix, const double *d_upot, const float2 *s1_in, float2 *s1_out)
{
    const int id = ix - P->nxfree;
    const bool dis = d_upot && id >= 1 && id <= P->nxdiso;
    if (dis) rg_launch_project(P, nb, id - 1, 0, d_upot, nb);
    const int slot = ix == 1 ? 0 : (ix == P->nx ? 2 : 1);
    if (schur_inversion(P, RG_MIXED_FP32)) {
        rg_launch_inverse_schur(P, nb, slot, ix == 1, dis, RG_MIXED_FP32);
    } else if (P->np >= 32) {
        if (inverse_kernel(P, RG_MIXED_FP32) == RG_INV_SYM) rg_launch_inverse_sym32(P, nb, slot, ix == 1, dis, P->d_redo);
        else rg_launch_inverse_nopiv(P, nb, slot, ix == 1, dis, P->d_redo);
        rg_launch_inverse_gj(P, nb, slot, ix == 1, dis, RG_MIXED_FP32, P->d_redo);
    } else {
        rg_launch_inverse_gj(P, nb, slot, ix == 1, dis, RG_MIXED_FP32);
    }
    if (ix == 1) rg_launch_copy_f2(P, P->d_gss32, s1_out, (size_t)nb * P->np * P->np);
    else rg_launch_gemm_sp(P, nb, P->d_gss32, s1_in, s1_out, RG_MIXED_FP32);
}

void fwd_step(rg_plan *P, int nb, int ix, const double *d_upot, const double2 *s1_in, double2 *s1_out)
{
    fwd_step_f64(P, nb, ix, d_upot, s1_in, s1_out);
}
void fwd_step(rg_plan *P, int nb, int ix, const double *d_upot, const float2 *s1_in, float2 *s1_out)
{
    fwd_step_f32(P, nb, ix, d_upot, s1_in, s1_out);
}
double2 *state_gss(rg_plan *P, double2 *) { return P->d_gss; }
float2 *state_gss(rg_plan *P, float2 *) { return P->d_gss32; }
void prod(rg_plan *P, int nb, double2 alpha, const double2 *A, int opA, const double2 *B, int opB, double2 beta, double2 *C,
          const double2 *dl, const double2 *dk, const double2 *dr)
{
    rg_zgemm(P, nb, alpha, A, opA, B, opB, beta, C, dl, dk, dr);
}
void prod(rg_plan *P, int nb, double2 alpha, const float2 *A, int opA, const float2 *B, int opB, double2 beta, float2 *C,
          const double2 *dl, const double2 *dk, const double2 *dr)
{
    rg_cgemm32(P, nb, alpha, A, opA, B, opB, beta, C, dl, dk, dr);
}
void topvec(rg_plan *P, int nb, const double2 *g, double2 *x) { rg_launch_tdt_topvec_f64(P, nb, g, P->d_status, x); }
void topvec(rg_plan *P, int nb, const float2 *g, double2 *x) { rg_launch_tdt_topvec_f32(P, nb, g, P->d_status, x); }

// Adjoint of L = Tr(t^H t) (or T_max) over nb realizations with explicit device potentials d_upot
// (layout (nxdiso, ny, nb)): the loss to d_loss[nb], dL/du to d_grad (same layout as d_upot).
// Forward sweep with a checkpoint (Gss, Gs1) after every adj_seg slices; then, segment by segment
// from the end, the states are recomputed from the checkpoint and the reverse recursion runs
// through them.  E = double2: complex128 state and products (FP64 MFMA); E = float2: complex64
// state (the mixed sweep's kernels) and products on FP32 MFMA, with the loss, the terminal adjoint
// and the gradient accumulated in FP64.
template <class E>
void run_adjoint(rg_plan *P, int nb, int loss_kind, const double *d_upot, double *d_loss, double *d_grad)
{
    const size_t mat = (size_t)nb * P->np * P->np, cmat = (size_t)P->adj_cap * P->np * P->np;
    const int nx = P->nx, S = P->adj_seg;
    const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
    E *ck = reinterpret_cast<E *>(P->d_ck), *sg = reinterpret_cast<E *>(P->d_seg), *ad = reinterpret_cast<E *>(P->d_adj);
    auto ck_g = [&](int c) { return ck + (size_t)(2 * c) * cmat; };
    auto ck_s = [&](int c) { return ck + (size_t)(2 * c + 1) * cmat; };
    auto seg_g = [&](int j) { return sg + (size_t)(2 * j) * cmat; };
    auto seg_s = [&](int j) { return sg + (size_t)(2 * j + 1) * cmat; };
    E *gbar = ad, *sbar = ad + cmat, *sbar2 = ad + 2 * cmat, *T1 = ad + 3 * cmat, *cbar = ad + 4 * cmat;
    E *gss = state_gss(P, (E *)nullptr);
    E *sa, *sb;
    if constexpr (std::is_same<E, double2>::value) {
        sa = P->d_gs1[0];
        sb = P->d_gs1[1];
    } else {
        sa = P->d_gs1_32[0];
        sb = P->d_gs1_32[1];
    }
    // forward sweep with checkpoints after slices S, 2S, ...
    for (int ix = 1; ix <= nx; ix++) {
        fwd_step(P, nb, ix, d_upot, sa, sb);
        std::swap(sa, sb);
        if (ix % S == 0 && ix < nx) {
            const int c = ix / S - 1;
            (void)hipMemcpyAsync(ck_g(c), gss, sizeof(E) * mat, hipMemcpyDeviceToDevice, P->stream);
            (void)hipMemcpyAsync(ck_s(c), sa, sizeof(E) * mat, hipMemcpyDeviceToDevice, P->stream);
        }
    }
    if (loss_kind == 0) {
        rg_launch_adj_terminal(P, nb, sa, d_loss, sbar);
    } else {  // T_max: top eigenvector of the Gram matrix (128 or 256 complex128 slots per realization in T1)
        double2 *xv = reinterpret_cast<double2 *>(T1);
        if (rg_topvec_supported(P)) {
            topvec(P, nb, sa, xv);
            rg_launch_adj_terminal_tmax(P, nb, sa, xv, 128, d_loss, sbar);
        } else {  // order 129..256: Jacobi with the eigenvectors
            const size_t wmat = (size_t)P->adj_cap * P->np * P->np;
            rg_launch_topvec_jacobi(P, nb, sa, P->d_adj_work, P->d_adj_work + wmat, xv, 256);
            rg_launch_adj_terminal_tmax(P, nb, sa, xv, 256, d_loss, sbar);
        }
    }
    (void)hipMemsetAsync(gbar, 0, sizeof(E) * mat, P->stream);
    (void)hipMemsetAsync(d_grad, 0, sizeof(double) * (size_t)nb * P->nxdiso * P->ny, P->stream);
    const int nseg = (nx + S - 1) / S;
    for (int c = nseg - 1; c >= 0; c--) {
        const int x0 = c * S + 1, x1 = std::min((c + 1) * S, nx);
        const E *s_in = nullptr;
        if (c >= 1) {
            (void)hipMemcpyAsync(gss, ck_g(c - 1), sizeof(E) * mat, hipMemcpyDeviceToDevice, P->stream);
            s_in = ck_s(c - 1);
        }
        for (int x = x0; x <= x1; x++) {  // recompute the segment's states
            const int j = x - x0;
            fwd_step(P, nb, x, d_upot, s_in, seg_s(j));
            (void)hipMemcpyAsync(seg_g(j), gss, sizeof(E) * mat, hipMemcpyDeviceToDevice, P->stream);
            s_in = seg_s(j);
        }
        for (int x = x1; x >= x0; x--) {  // reverse recursion
            const int j = x - x0;
            const E *gx = seg_g(j);
            if (x >= 2) {
                const E *sprev = j > 0 ? seg_s(j - 1) : ck_s(c - 1);
                prod(P, nb, mone, sbar, 0, sprev, 1, one, gbar, nullptr, nullptr, P->d_bconj);
                prod(P, nb, mone, gx, 1, sbar, 0, zero, sbar2, P->d_bconj, nullptr, nullptr);
                std::swap(sbar, sbar2);
            } else {
                rg_launch_adj_add(P, nb, gbar, sbar);
            }
            prod(P, nb, one, gbar, 0, gx, 1, zero, T1, nullptr, nullptr, nullptr);
            prod(P, nb, mone, gx, 1, T1, 0, zero, cbar, nullptr, nullptr, nullptr);
            const int id = x - P->nxfree;
            if (id >= 1 && id <= P->nxdiso) rg_launch_adj_grad(P, nb, cbar, id - 1, d_grad);
            if (x >= 2) rg_launch_adj_scale(P, nb, gbar, cbar);
        }
    }
}

void run_adjoint_any(rg_plan *P, int nb, int loss_kind, const double *d_upot, double *d_loss, double *d_grad)
{
    if (P->adj_prec == RG_MIXED_FP32) run_adjoint<float2>(P, nb, loss_kind, d_upot, d_loss, d_grad);
    else run_adjoint<double2>(P, nb, loss_kind, d_upot, d_loss, d_grad);
}

// Streams, events and plan views of the two lanes (created once, on first use).
rg_status ensure_lanes(rg_plan *P)
{
    if (P->lane[0]) return RG_OK;
    for (hipEvent_t &e : P->lane_ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return RG_EDEVICE;
    for (rg_plan *&L : P->lane) {
        L = new (std::nothrow) rg_plan(*P);
        if (!L) return RG_ENOMEM;
        L->lanes = 1;
        L->lane[0] = L->lane[1] = nullptr;
        for (hipEvent_t &e : L->lane_ev) e = nullptr;
        for (auto &t : L->timers) { t.pending.clear(); t.pool.clear(); t.ms = 0.0; t.launches = 0; }
        L->graphs.clear();  // graphs belong to the plan; lanes launch directly
        L->use_graph = false;
        L->stream = nullptr;
        if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess) return RG_EDEVICE;
    }
    return RG_OK;
}

int inverse_choice(const char *v, int dflt, bool fp64)
{
    if (!v) return dflt;
    if (fp64 && std::strcmp(v, "register") == 0) return RG_INV_REGISTER;
    if (std::strcmp(v, "lookahead") == 0) return RG_INV_LOOKAHEAD;
    if (std::strcmp(v, "blocked") == 0) return RG_INV_BLOCKED;
    if (std::strcmp(v, "verified") == 0) return RG_INV_NOPIV;
    if (std::strcmp(v, "symmetric") == 0) return RG_INV_SYM;
    return dflt;
}

rg_status check_common(rg_plan *P, int32_t n, int32_t precision)
{
    if (!P || n < 0) return RG_EINVAL;
    if (precision != RG_FP64 && precision != RG_MIXED_FP32 && precision != RG_MIXED_BF16X3) return RG_EINVAL;
    if (P->np > 256) return RG_ENOTSUP;
    return RG_OK;
}

rg_status harvest_status(rg_plan *P, int nb, int32_t *status_host, int32_t *worst)
{
    std::vector<int32_t> st(nb);
    HIPCHK(hipMemcpyAsync(st.data(), P->d_status, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    for (int i = 0; i < nb; i++) {
        if (status_host) status_host[i] = st[i];
        if (st[i] && !*worst) *worst = st[i];
    }
    return RG_OK;
}

// ------------------------------------------------------------------------------------------------
// Uneven batches: realizations with their own length, absorption, disorder strength and seed.
// ------------------------------------------------------------------------------------------------
struct uneven_real {
    int idx;  // position in the caller's arrays
    int nx, nxdiso;
    double stdu;
    int64_t seed;
    std::vector<double2> adiag, b, sqrtmu;  // 3 np (first, bulk, last slice), np, np
};

// One group (same boundary, mode count, cells, filter ranges and free length) through plan P:
// realizations sorted by slice count, longest first, in chunks of P->batch_max.  Slice ix runs
// over the prefix of realizations still active (nx >= ix) -- one launch per kernel, each
// realization taking its last diagonal on its own last slice (nx_r) -- split only where the
// disorder region ends (free slices).  Realizations whose last slice it was have their G(L;0)
// copied into their own (now unused) Gss slot; one eigenvalue launch per chunk follows the sweep.
// HIP error inside run_uneven_group: free the group's device buffers before returning.
#define UCHK(x)                                       \
    do {                                              \
        if ((x) != hipSuccess) {                      \
            release();                                \
            return RG_EDEVICE;                        \
        }                                             \
    } while (0)
rg_status run_uneven_group(rg_plan *P, std::vector<uneven_real> &rs, int precision, double *tallvals,
                           const int64_t *offsets, int32_t *status, int32_t *worst)
{
    const int np = P->np, neigv = P->neigv, cap = P->batch_max;
    std::stable_sort(rs.begin(), rs.end(), [](const uneven_real &a, const uneven_real &b) { return a.nx > b.nx; });
    double2 *d_ad = nullptr, *d_b = nullptr, *d_sm = nullptr;
    int64_t *d_seeds = nullptr;
    double *d_sd = nullptr;
    int32_t *d_nxr = nullptr;
    auto release = [&]() {
        for (void *q : {(void *)d_ad, (void *)d_b, (void *)d_sm, (void *)d_seeds, (void *)d_sd, (void *)d_nxr})
            if (q) (void)hipFree(q);
    };
    rg_status st;
    if ((st = dalloc(&d_ad, (size_t)cap * 3 * np)) || (st = dalloc(&d_b, (size_t)cap * np)) ||
        (st = dalloc(&d_sm, (size_t)cap * np)) || (st = dalloc(&d_seeds, (size_t)cap)) || (st = dalloc(&d_sd, (size_t)cap)) ||
        (st = dalloc(&d_nxr, (size_t)cap))) {
        release();
        return st;
    }
    // base view: the plan's buffers with per-realization constants; seg: a range of it
    rg_plan V(*P);
    V.lanes = 1;
    V.lane[0] = V.lane[1] = nullptr;
    V.graphs.clear();
    V.use_graph = false;
    V.profiling = false;
    for (auto &t : V.timers) { t.pending.clear(); t.pool.clear(); }
    V.cstride = np;
    V.d_adiag = d_ad;
    V.d_b = d_b;
    V.d_sqrtmu = d_sm;
    V.d_seeds_r = d_seeds;
    V.d_stdu_r = d_sd;
    V.d_nx_r = d_nxr;
    rg_plan S(V);
    const size_t mat = (size_t)np * np;
    std::vector<double2> had((size_t)cap * 3 * np), hb((size_t)cap * np), hsm((size_t)cap * np);
    std::vector<int64_t> hseed(cap);
    std::vector<double> hsd(cap), te((size_t)cap * neigv);
    std::vector<int32_t> sv(cap), hnx(cap);
    for (size_t c0 = 0; c0 < rs.size(); c0 += cap) {
        const int nb = (int)std::min<size_t>(cap, rs.size() - c0);
        for (int i = 0; i < nb; i++) {
            const uneven_real &r = rs[c0 + i];
            std::copy(r.adiag.begin(), r.adiag.end(), had.begin() + (size_t)i * 3 * np);
            std::copy(r.b.begin(), r.b.end(), hb.begin() + (size_t)i * np);
            std::copy(r.sqrtmu.begin(), r.sqrtmu.end(), hsm.begin() + (size_t)i * np);
            hseed[i] = r.seed;
            hsd[i] = r.stdu;
            hnx[i] = r.nx;
        }
        UCHK(hipMemcpy(d_ad, had.data(), sizeof(double2) * (size_t)nb * 3 * np, hipMemcpyHostToDevice));
        UCHK(hipMemcpy(d_b, hb.data(), sizeof(double2) * (size_t)nb * np, hipMemcpyHostToDevice));
        UCHK(hipMemcpy(d_sm, hsm.data(), sizeof(double2) * (size_t)nb * np, hipMemcpyHostToDevice));
        UCHK(hipMemcpy(d_seeds, hseed.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice));
        UCHK(hipMemcpy(d_sd, hsd.data(), sizeof(double) * nb, hipMemcpyHostToDevice));
        UCHK(hipMemcpy(d_nxr, hnx.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice));
        UCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        int cur = 0;
        const int nxmax = rs[c0].nx;
        for (int ix = 1; ix <= nxmax; ix++) {
            int nact = nb;
            while (nact > 0 && rs[c0 + nact - 1].nx < ix) nact--;
            V.cur_ix = ix;
            auto ixd_of = [&](int i) {
                const int id = ix - P->nxfree;
                return id >= 1 && id <= rs[c0 + i].nxdiso ? id - 1 : -1;
            };
            for (int a = 0; a < nact;) {
                const int k = ixd_of(a);
                int e = a + 1;
                while (e < nact && ixd_of(e) == k) e++;
                lane_view(&V, &S, a, e - a);
                sweep_step(&S, e - a, ix, ix == 1 ? 0 : 1, k, 0, nullptr, precision, &cur, false);
                a = e;
            }
            if (ix > 1) cur ^= 1;
            int f0 = nact;
            while (f0 > 0 && rs[c0 + f0 - 1].nx == ix) f0--;
            if (f0 < nact) {  // G(L;0) of the realizations that ended here, kept in their own Gss slot
                // (no later slice touches a finished realization's slot; the t buffer stays free
                // as the work area of the Jacobi singular values)
                const size_t n2 = (size_t)(nact - f0) * mat, o = (size_t)f0 * mat;
                if (precision == RG_FP64)
                    rg_launch_copy_f2(P, reinterpret_cast<const float2 *>(P->d_gs1[cur] + o),
                                      reinterpret_cast<float2 *>(P->d_gss + o), 2 * n2);
                else
                    rg_launch_copy_f2(P, P->d_gs1_32[cur] + o, P->d_gss32 + o, n2);
            }
        }
        if (precision == RG_FP64) rg_launch_teig_f64(&V, nb, P->d_gss, P->d_teig, P->d_status);
        else rg_launch_teig_f32(&V, nb, P->d_gss32, P->d_teig, P->d_status);
        UCHK(hipGetLastError());
        UCHK(hipMemcpyAsync(te.data(), P->d_teig, sizeof(double) * (size_t)nb * neigv, hipMemcpyDeviceToHost, P->stream));
        UCHK(hipMemcpyAsync(sv.data(), P->d_status, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, P->stream));
        UCHK(hipStreamSynchronize(P->stream));
        for (int i = 0; i < nb; i++) {
            const int idx = rs[c0 + i].idx;
            if (status) status[idx] = sv[i];
            if (sv[i] && !*worst) *worst = sv[i];
            std::copy(te.begin() + (size_t)i * neigv, te.begin() + (size_t)(i + 1) * neigv, tallvals + offsets[idx]);
        }
    }
    release();
    return RG_OK;
}
#undef UCHK

}  // namespace

// Method of the transmission eigenvalues for a state precision (option eigen).
static int eigen_method(const rg_plan *P, bool fp64)
{
    int e = P->eigen == RG_EIG_AUTO ? (fp64 ? RG_EIG_BIDIAG : RG_EIG_GRAM) : P->eigen;
    if (e == RG_EIG_BIDIAG && !rg_tsvd_supported(P)) e = RG_EIG_JACOBI;
    if (e == RG_EIG_GRAM && !rg_eig_supported(P)) e = RG_EIG_JACOBI;
    return e;
}

void rg_launch_teig_f64(rg_plan *P, int nb, const double2 *gs1, double *teig, int32_t *status)
{
    switch (eigen_method(P, true)) {
        case RG_EIG_BIDIAG: rg_launch_tsvd_f64(P, nb, gs1, teig, status); break;
        case RG_EIG_GRAM: rg_launch_tdt_eig_f64(P, nb, gs1, teig, status); break;
        default: rg_launch_fisher_lee_svd_f64(P, nb, gs1, teig, status, nullptr);
    }
}

void rg_launch_teig_f32(rg_plan *P, int nb, const float2 *gs1, double *teig, int32_t *status)
{
    switch (eigen_method(P, false)) {
        case RG_EIG_BIDIAG: rg_launch_tsvd_f32(P, nb, gs1, teig, status); break;
        case RG_EIG_GRAM: rg_launch_tdt_eig_f32(P, nb, gs1, teig, status); break;
        default: rg_launch_fisher_lee_svd_f32(P, nb, gs1, teig, status, nullptr);
    }
}


// ------------------------------------------------------------------------------------------------
// Timing.
// ------------------------------------------------------------------------------------------------
static const char *kTimerNames[RGK_COUNT] = {"project", "slice_inverse", "dyson_gemm", "fisher_lee_svd",
                                             "histogram", "copy", "inverse_redo"};

void rg_timer_begin(rg_plan *P, int id, hipEvent_t *stop_out)
{
    *stop_out = nullptr;
    if (!P->profiling || !((P->profile_mask >> id) & 1)) return;
    rg_timer &t = P->timers[id];
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (!t.pool.empty()) {
        ev = t.pool.back();
        t.pool.pop_back();
    } else {
        if (hipEventCreate(&ev.first) != hipSuccess) return;
        if (hipEventCreate(&ev.second) != hipSuccess) { (void)hipEventDestroy(ev.first); return; }
    }
    (void)hipEventRecord(ev.first, P->stream);
    t.pending.push_back(ev);
    *stop_out = ev.second;
}

void rg_timer_end(rg_plan *P, int id, hipEvent_t stop)
{
    if (!P->profiling || !stop) return;
    (void)hipEventRecord(stop, P->stream);
    (void)id;
}

static void harvest_one(rg_timer &t)
{
    for (auto &e : t.pending) {
        float ms = 0.f;
        if (hipEventSynchronize(e.second) == hipSuccess && hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
            t.ms += ms;
            t.launches++;
        }
        t.pool.push_back(e);
    }
    t.pending.clear();
}

// Launch times of the plan's stream and of its lanes' streams, added into the plan's timers.
static void harvest_timers(rg_plan *P)
{
    for (auto &t : P->timers) harvest_one(t);
    for (rg_plan *L : P->lane) {
        if (!L) continue;
        for (int i = 0; i < RGK_COUNT; i++) {
            harvest_one(L->timers[i]);
            P->timers[i].ms += L->timers[i].ms;
            P->timers[i].launches += L->timers[i].launches;
            L->timers[i].ms = 0.0;
            L->timers[i].launches = 0;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// C ABI.
// ------------------------------------------------------------------------------------------------
extern "C" {

rg_status rg_plan_create(const rg_params *params, int32_t batch_max, int32_t device, rg_plan **plan)
{
    if (!params || !plan || batch_max < 1) return RG_EINVAL;
    *plan = nullptr;
    rg_plan *P = new (std::nothrow) rg_plan();
    if (!P) return RG_ENOMEM;
    P->p = *params;
    P->device = device;
    P->batch_max = batch_max;
    for (int i = 0; i < RGK_COUNT; i++) P->timers[i].name = kTimerNames[i];
    // Slice-inversion kernels, overridable for measurement: RG_INVERSE_FP64 and RG_INVERSE_MIXED in
    // {lookahead, blocked, verified}, RG_INVERSE_FP64 also register and symmetric.
    P->inv_fp64 = inverse_choice(std::getenv("RG_INVERSE_FP64"), P->inv_fp64, true);
    P->inv_mixed = inverse_choice(std::getenv("RG_INVERSE_MIXED"), P->inv_mixed, false);
    // Transmission eigenvalues: RG_EIGEN in {gram, jacobi, bidiag} overrides the default (bidiagonal SVD
    // of t for complex128 sweeps, eigenvalues of t^H t for complex64 sweeps).
    const char *eg = std::getenv("RG_EIGEN");
    if (eg && std::strcmp(eg, "gram") == 0) P->eigen = RG_EIG_GRAM;
    else if (eg && std::strcmp(eg, "jacobi") == 0) P->eigen = RG_EIG_JACOBI;
    else if (eg && std::strcmp(eg, "bidiag") == 0) P->eigen = RG_EIG_BIDIAG;
    // complex64 Dyson product: RG_DYSON=tile64 for the 64 x 64 tiled kernel, RG_DYSON=bf16x6 for the
    // BF16 matrix-core kernel, RG_DYSON=gauss3 for three real products per complex product, instead of
    // one workgroup per realization on FP32 MFMA with four.
    const char *dy = std::getenv("RG_DYSON");
    P->gemm_wg = dy && std::strcmp(dy, "tile64") == 0 ? 0
                 : dy && std::strcmp(dy, "bf16x6") == 0 ? 2
                 : dy && std::strcmp(dy, "gauss3") == 0 ? 3
                                                        : P->gemm_wg;
    // Projection kernels: RG_PROJECT_WIDE_MIN = the smallest batch of the wide / generating kernels (default 1:
    // every batch; larger values send smaller batches to k_project_mfma, for measurement).
    if (const char *pw = std::getenv("RG_PROJECT_WIDE_MIN")) P->project_wide_min = std::max(1, std::atoi(pw));
    rg_status st = init_modes(P);  // argument checks first, like init_waveguide
    if (st) { delete P; return st; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) { delete P; return RG_EDEVICE; }
    DeviceGuard g(device);
    if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess) { free_plan(P); return RG_EDEVICE; }
    if (hipDeviceGetAttribute(&P->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || P->ncu < 1) {
        free_plan(P);
        return RG_EDEVICE;
    }
    if ((st = upload_constants(P)) || (st = alloc_batch(P))) { free_plan(P); return st; }
    // Concurrent lanes (RG_LANES=1 or 2): views of this plan on streams of their own.
    const char *ln = std::getenv("RG_LANES");
    const int lanes = ln ? std::max(1, std::min(2, std::atoi(ln))) : P->lanes;
    if (lanes > 1 && (st = ensure_lanes(P))) { free_plan(P); return st; }
    P->lanes = lanes;
    *plan = P;
    return RG_OK;
}

rg_status rg_plan_destroy(rg_plan *plan)
{
    if (!plan) return RG_EINVAL;
    (void)hipStreamSynchronize(plan->stream);
    free_plan(plan);
    return RG_OK;
}

rg_status rg_plan_dims(const rg_plan *P, int32_t *dims, double *vals)
{
    if (!P || !dims) return RG_EINVAL;
    const int d[11] = {P->nmu, P->ny, P->nx, P->neigv, P->imin, P->imax, P->jmin, P->jmax, P->nprop, P->nxdiso, P->nxfree};
    for (int i = 0; i < 11; i++) dims[i] = d[i];
    if (vals) { vals[0] = P->kl; vals[1] = P->kdx; vals[2] = P->dos; vals[3] = P->stdu; }
    return RG_OK;
}

rg_status rg_tdistrib_device(rg_plan *P, int64_t seed0, int32_t nseed, int32_t precision, double *tallvals_dev,
                             int32_t *status_dev)
{
    rg_status st = check_common(P, nseed, precision);
    if (st) return st;
    if (!tallvals_dev && nseed > 0) return RG_EINVAL;
    DeviceGuard g(P->device);
    int cap = 0;
    if ((st = pass_cap(P, precision, P->batch_max, &cap))) return st;
    for (int32_t off = 0; off < nseed; off += cap) {
        const int nb = std::min(cap, nseed - off);
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        run_sweep(P, nb, seed0 + off, nullptr, precision, tallvals_dev + (size_t)off * P->neigv, nullptr);
        if (status_dev)
            HIPCHK(hipMemcpyAsync(status_dev + off, P->d_status, sizeof(int32_t) * nb, hipMemcpyDeviceToDevice, P->stream));
    }
    HIPCHK(hipGetLastError());
    return RG_OK;
}

rg_status rg_tdistrib(rg_plan *P, int64_t seed0, int32_t nseed, int32_t precision, double *tallvals, int32_t *status)
{
    rg_status st = check_common(P, nseed, precision);
    if (st) return st;
    if (!tallvals && nseed > 0) return RG_EINVAL;
    DeviceGuard g(P->device);
    int32_t worst = 0, cap = 0;
    if ((st = pass_cap(P, precision, P->batch_max, &cap))) return st;
    for (int32_t off = 0; off < nseed; off += cap) {
        const int nb = std::min(cap, nseed - off);
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        run_sweep(P, nb, seed0 + off, nullptr, precision, P->d_teig, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(tallvals + (size_t)off * P->neigv, P->d_teig, sizeof(double) * (size_t)nb * P->neigv,
                              hipMemcpyDeviceToHost, P->stream));
        if ((st = harvest_status(P, nb, status ? status + off : nullptr, &worst))) return st;
    }
    if (P->profiling) harvest_timers(P);
    return worst ? (rg_status)worst : RG_OK;
}

rg_status rg_tdistrib_uneven(const rg_params *params, const int64_t *seeds, int32_t nreal, int32_t precision,
                             int32_t device, int32_t batch_max, double *tallvals, int64_t *offsets, int32_t *status)
{
    if (!params || !seeds || !offsets || nreal < 0 || batch_max < 1) return RG_EINVAL;
    if (precision != RG_FP64 && precision != RG_MIXED_FP32 && precision != RG_MIXED_BF16X3) return RG_EINVAL;
    // per-realization constants on the host; groups of realizations the kernels can batch
    std::map<std::array<int, 9>, std::vector<uneven_real>> groups;
    int32_t worst = 0;
    offsets[0] = 0;
    for (int32_t r = 0; r < nreal; r++) {
        rg_plan t;
        t.p = params[r];
        const rg_status st = init_modes(&t);
        offsets[r + 1] = offsets[r] + (st ? 0 : t.neigv);
        if (st) {
            if (status) status[r] = st;
            if (!worst) worst = st;
            continue;
        }
        uneven_real u;
        u.idx = r;
        u.nx = t.nx;
        u.nxdiso = t.nxdiso;
        u.stdu = t.stdu;
        u.seed = seeds[r];
        u.adiag.assign(3 * (size_t)t.np, make_double2(0, 0));
        for (int m = 0; m < t.nmu; m++) {  // as upload_constants
            u.adiag[m] = t.h_abound[m];
            u.adiag[t.np + m] = t.h_abulk[m];
            u.adiag[2 * t.np + m] = t.nx == 2 ? cadd(t.h_abound[m], t.h_abulk[m]) : t.h_abound[m];
        }
        u.b = t.h_b;
        u.sqrtmu = t.h_sqrtmu;
        groups[{t.p.boundtype, t.nmu, t.ny, t.imin, t.imax, t.jmin, t.jmax, t.neigv, t.nxfree}].push_back(std::move(u));
    }
    if (!tallvals && offsets[nreal] > 0) return RG_EINVAL;
    // The groups run at the same time, each from a host thread of its own on its own plan and stream: a
    // group whose launches do not fill the device (few or short realizations, or the tail of its sweep
    // where only the longest remain) leaves room to the others.  Every realization meets the same kernels
    // and data as in a group of its own, so the results do not depend on this schedule.
    // A large group above 128 modes (the most costly per slice) is swept as two halves on two streams,
    // its realizations dealt alternately by length: the halves fill each other's partly empty last
    // rounds of workgroups and the gaps between the short kernels of a slice step (10 000 realizations
    // with N in {16, ..., 256}, L/l_s in [0.5, 10] and absorption: FP64 14.3 -> 13.5 s, mixed 6.1 -> 6.0 s;
    // three or four parts gained nothing more).
    std::vector<std::vector<uneven_real> *> gl;
    std::list<std::vector<uneven_real>> halves;
    for (auto &kv : groups) {
        std::vector<uneven_real> &rs = kv.second;
        if (rs[0].b.size() > 128 && rs[0].b.size() <= 256 && rs.size() >= 512) {
            std::stable_sort(rs.begin(), rs.end(), [](const uneven_real &a, const uneven_real &b) { return a.nx > b.nx; });
            halves.emplace_back();
            halves.emplace_back();
            std::vector<uneven_real> &h0 = *std::prev(halves.end(), 2), &h1 = halves.back();
            for (size_t i = 0; i < rs.size(); i++) (i % 2 ? h1 : h0).push_back(std::move(rs[i]));
            gl.push_back(&h0);
            gl.push_back(&h1);
        } else {
            gl.push_back(&rs);
        }
    }
    // The plans (and their streams) are created here, costliest group first (cost: modes^3 x slices),
    // whatever the order in which the host threads start: the streams of a process share the device's
    // hardware queues in creation order, and a fixed order makes the schedule, and the time, the same from
    // run to run (uneven batch of 10 000 realizations with N in {16, ..., 256}: FP64 13.2 s and mixed 6.2 s
    // on three runs, against 13.4-14.3 s and 5.9-6.7 s with the plans created by the threads; creating the
    // cheapest groups first gave 15.3 s and 6.0 s).
    std::vector<double> cost(gl.size(), 0.0);
    for (size_t k = 0; k < gl.size(); k++) {
        const double n = (double)(*gl[k])[0].b.size();
        for (const uneven_real &u : *gl[k]) cost[k] += n * n * n * u.nx;
    }
    std::vector<size_t> order(gl.size());
    for (size_t k = 0; k < gl.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<int32_t> gworst(gl.size(), 0);
    std::vector<rg_plan *> plans(gl.size(), nullptr);
    std::vector<rg_status> pst(gl.size(), RG_OK);
    for (size_t k : order) {
        std::vector<uneven_real> &rs = *gl[k];
        const int cap = (int)std::min<size_t>((size_t)batch_max, rs.size());
        // mode counts above 256 are refused before any device memory is taken
        pst[k] = rs[0].b.size() > 256 ? RG_ENOTSUP : rg_plan_create(&params[rs[0].idx], cap, device, &plans[k]);
    }
    auto run_group = [&](size_t k) {
        std::vector<uneven_real> &rs = *gl[k];
        rg_plan *P = plans[k];
        rg_status st = pst[k];
        if (!st) {
            DeviceGuard g(P->device);
            st = run_uneven_group(P, rs, precision, tallvals, offsets, status, &gworst[k]);
        }
        if (st) {  // the whole group: set-up or device failure
            for (const uneven_real &u : rs)
                if (status) status[u.idx] = st;
            if (!gworst[k]) gworst[k] = st;
        }
    };
    std::vector<std::thread> threads;
    size_t k0 = 1;
    try {
        for (; k0 < order.size(); k0++) threads.emplace_back(run_group, order[k0]);
    } catch (...) {  // no thread available: the remaining groups run on this one
    }
    if (!order.empty()) run_group(order[0]);
    for (size_t k = k0; k < order.size(); k++) run_group(order[k]);
    for (auto &t : threads) t.join();
    for (rg_plan *P : plans)
        if (P) rg_plan_destroy(P);
    for (int32_t w : gworst)
        if (!worst && w) worst = w;
    return worst ? (rg_status)worst : RG_OK;
}

static rg_status potential_pass(rg_plan *P, const double *upot, int32_t nreal, int32_t precision, double *out,
                                int32_t *status, bool want_t)
{
    rg_status st = check_common(P, nreal, precision);
    if (st) return st;
    if (!out && nreal > 0) return RG_EINVAL;
    DeviceGuard g(P->device);
    const size_t per = (size_t)P->nxdiso * P->ny;
    const int m = P->imax - P->imin + 1, nn = P->jmax - P->jmin + 1;
    int cap = 0;
    if ((st = pass_cap(P, precision, std::min(P->upot_cap, P->batch_max), &cap))) return st;
    int32_t worst = 0;
    for (int32_t off = 0; off < nreal; off += cap) {
        const int nb = std::min(cap, nreal - off);
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        const double *dup = nullptr;
        if (upot) {
            for (size_t e = 0; e < per * nb; e++)
                if (!std::isfinite(upot[per * off + e])) return RG_ENONFINITE;
            HIPCHK(hipMemcpyAsync(P->d_upot, upot + per * off, sizeof(double) * per * nb, hipMemcpyHostToDevice, P->stream));
            dup = P->d_upot;
        } else {
            HIPCHK(hipMemsetAsync(P->d_upot, 0, sizeof(double) * per * nb, P->stream));
            dup = P->d_upot;
        }
        if (want_t) {
            run_sweep(P, nb, 0, dup, precision, nullptr, P->d_tmat);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(out + (size_t)off * m * nn * 2, P->d_tmat, sizeof(double2) * (size_t)nb * m * nn,
                                  hipMemcpyDeviceToHost, P->stream));
        } else {
            run_sweep(P, nb, 0, dup, precision, P->d_teig, nullptr);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(out + (size_t)off * P->neigv, P->d_teig, sizeof(double) * (size_t)nb * P->neigv,
                                  hipMemcpyDeviceToHost, P->stream));
        }
        if ((st = harvest_status(P, nb, status ? status + off : nullptr, &worst))) return st;
    }
    if (P->profiling) harvest_timers(P);
    return worst ? (rg_status)worst : RG_OK;
}

rg_status rg_tdistrib_potential(rg_plan *P, const double *upot, int32_t nreal, int32_t precision, double *tallvals,
                                int32_t *status)
{
    if (!upot && nreal > 0) return RG_EINVAL;
    return potential_pass(P, upot, nreal, precision, tallvals, status, false);
}

rg_status rg_transmission_potential(rg_plan *P, const double *upot, int32_t nreal, int32_t precision, double *tmat,
                                    int32_t *status)
{
    return potential_pass(P, upot, nreal, precision, tmat, status, true);
}

// Reflection and transmission matrices, FP64 (compute_refl_trans_v1, src/waveguide.f90:469-519),
// for seeds (upot == nullptr) or explicit host potentials.
static rg_status refl_trans_pass(rg_plan *P, int64_t seed0, const double *upot, int32_t nreal, double *rmat,
                                 double *tmat, int32_t *status)
{
    rg_status st = check_common(P, nreal, RG_FP64);
    if (st) return st;
    if ((!rmat || !tmat) && nreal > 0) return RG_EINVAL;
    DeviceGuard g(P->device);
    const size_t per = (size_t)P->nxdiso * P->ny, nn2 = (size_t)P->nmu * P->nmu;
    const int cap = upot ? std::min(P->upot_cap, P->refl_cap) : P->refl_cap;
    double2 *d_r = P->d_tmat;  // batch_max x np^2 complex >= nb x nmu^2
    int32_t worst = 0;
    for (int32_t off = 0; off < nreal; off += cap) {
        const int nb = std::min(cap, nreal - off);
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        const double *dup = nullptr;
        if (upot) {
            for (size_t e = 0; e < per * nb; e++)
                if (!std::isfinite(upot[per * off + e])) return RG_ENONFINITE;
            HIPCHK(hipMemcpyAsync(P->d_upot, upot + per * off, sizeof(double) * per * nb, hipMemcpyHostToDevice, P->stream));
            dup = P->d_upot;
        }
        const double2 *d_t = run_sweep_refl(P, nb, seed0 + off, dup, d_r);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(rmat + 2 * nn2 * off, d_r, sizeof(double2) * nn2 * nb, hipMemcpyDeviceToHost, P->stream));
        HIPCHK(hipMemcpyAsync(tmat + 2 * nn2 * off, d_t, sizeof(double2) * nn2 * nb, hipMemcpyDeviceToHost, P->stream));
        if ((st = harvest_status(P, nb, status ? status + off : nullptr, &worst))) return st;
    }
    if (P->profiling) harvest_timers(P);
    return worst ? (rg_status)worst : RG_OK;
}

rg_status rg_refl_trans(rg_plan *P, int64_t seed0, int32_t nseed, double *rmat, double *tmat, int32_t *status)
{
    return refl_trans_pass(P, seed0, nullptr, nseed, rmat, tmat, status);
}

rg_status rg_refl_trans_potential(rg_plan *P, const double *upot, int32_t nreal, double *rmat, double *tmat,
                                  int32_t *status)
{
    if (!upot && nreal > 0) return RG_EINVAL;
    return refl_trans_pass(P, 0, upot, nreal, rmat, tmat, status);
}

static rg_status deposition_pass(rg_plan *P, int64_t seed0, const double *upot, int32_t nreal, int32_t fpos,
                                 int32_t ipos, double *depmat, int32_t *status)
{
    rg_status st = check_common(P, nreal, RG_FP64);
    if (st) return st;
    if ((!depmat && nreal > 0) || fpos < 1 || fpos > P->nx || ipos < 1 || ipos > P->nx) return RG_EINVAL;
    DeviceGuard g(P->device);
    const size_t per = (size_t)P->nxdiso * P->ny, nn2 = (size_t)P->nmu * P->nmu;
    const int cap = upot ? std::min(P->upot_cap, P->refl_cap) : P->refl_cap;
    double2 *d_dep = P->d_tmat;  // batch_max x np^2 complex >= nb x nmu^2
    int32_t worst = 0;
    for (int32_t off = 0; off < nreal; off += cap) {
        const int nb = std::min(cap, nreal - off);
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        const double *dup = nullptr;
        if (upot) {
            for (size_t e = 0; e < per * nb; e++)
                if (!std::isfinite(upot[per * off + e])) return RG_ENONFINITE;
            HIPCHK(hipMemcpyAsync(P->d_upot, upot + per * off, sizeof(double) * per * nb, hipMemcpyHostToDevice, P->stream));
            dup = P->d_upot;
        }
        run_sweep_deposition(P, nb, seed0 + off, dup, fpos, ipos, d_dep);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(depmat + 2 * nn2 * off, d_dep, sizeof(double2) * nn2 * nb, hipMemcpyDeviceToHost, P->stream));
        if ((st = harvest_status(P, nb, status ? status + off : nullptr, &worst))) return st;
    }
    if (P->profiling) harvest_timers(P);
    return worst ? (rg_status)worst : RG_OK;
}

rg_status rg_deposition(rg_plan *P, int64_t seed0, int32_t nseed, int32_t fpos, int32_t ipos, double *depmat,
                        int32_t *status)
{
    return deposition_pass(P, seed0, nullptr, nseed, fpos, ipos, depmat, status);
}

rg_status rg_deposition_potential(rg_plan *P, const double *upot, int32_t nreal, int32_t fpos, int32_t ipos,
                                  double *depmat, int32_t *status)
{
    if (!upot && nreal > 0) return RG_EINVAL;
    return deposition_pass(P, 0, upot, nreal, fpos, ipos, depmat, status);
}

rg_status rg_tprofile_reserve(rg_plan *P, int32_t nreal_max)
{
    if (!P || nreal_max < 1) return RG_EINVAL;
    if (!rg_tprofile_supported(P->nmu)) return RG_ENOTSUP;
    DeviceGuard g(P->device);
    if (P->tp_cap >= nreal_max) return RG_OK;
    for (void **b : {(void **)&P->d_tp_g, (void **)&P->d_tp_z, (void **)&P->d_tp_x, (void **)&P->d_tp_v,
                     (void **)&P->d_tp_w, (void **)&P->d_tp_y[0], (void **)&P->d_tp_y[1], (void **)&P->d_tp_t,
                     (void **)&P->d_tp_i, (void **)&P->d_tp_psi}) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    P->tp_cap = 0;
    const size_t cmat = (size_t)nreal_max * P->np * P->np;
    rg_status st;
    if ((st = dalloc(&P->d_tp_g, (size_t)P->nx * cmat)) || (st = dalloc(&P->d_tp_z, (size_t)P->nx * cmat)) ||
        (st = dalloc(&P->d_tp_x, cmat)) || (st = dalloc(&P->d_tp_v, cmat)) || (st = dalloc(&P->d_tp_w, cmat)) ||
        (st = dalloc(&P->d_tp_y[0], cmat)) || (st = dalloc(&P->d_tp_y[1], cmat)) ||
        (st = dalloc(&P->d_tp_t, (size_t)nreal_max * P->nmu)) ||
        (st = dalloc(&P->d_tp_i, (size_t)nreal_max * P->nx * P->nmu)) ||
        (st = dalloc(&P->d_tp_psi, (size_t)P->nx * P->ny)))
        return st;
    P->tp_cap = nreal_max;
    return RG_OK;
}

static rg_status tprofile_pass(rg_plan *P, int64_t seed0, const double *upot, int32_t nreal, int32_t filtered,
                               double *tvals, double *intensity, int32_t *status)
{
    rg_status st = check_common(P, nreal, RG_FP64);
    if (st) return st;
    if ((!tvals || !intensity) && nreal > 0) return RG_EINVAL;
    if (P->tp_cap < 1) return RG_EINVAL;
    // the group variant decomposes the aperture-filtered t; it needs no more columns than rows
    // (diagonalize_ahca's size check, src/base_utils.f90:411-415)
    if (filtered && P->jmax - P->jmin > P->imax - P->imin) return RG_EINVAL;
    DeviceGuard g(P->device);
    const size_t nvr = filtered ? (size_t)(P->jmax - P->jmin + 1) : (size_t)P->nmu;
    const size_t per = (size_t)P->nxdiso * P->ny, n = nvr, nxn = (size_t)P->nx * nvr;
    // a pass also uses the batch buffers (state, status, redo flags) of batch_max realizations
    const int cap = std::min(upot ? std::min(P->upot_cap, P->tp_cap) : P->tp_cap, P->batch_max);
    int32_t worst = 0;
    for (int32_t off = 0; off < nreal; off += cap) {
        const int nb = std::min(cap, nreal - off);
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        const double *dup = nullptr;
        if (upot) {
            for (size_t e = 0; e < per * nb; e++)
                if (!std::isfinite(upot[per * off + e])) return RG_ENONFINITE;
            HIPCHK(hipMemcpyAsync(P->d_upot, upot + per * off, sizeof(double) * per * nb, hipMemcpyHostToDevice, P->stream));
            dup = P->d_upot;
        }
        run_sweep_tprofile(P, nb, seed0 + off, dup, filtered != 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(tvals + n * off, P->d_tp_t, sizeof(double) * n * nb, hipMemcpyDeviceToHost, P->stream));
        HIPCHK(hipMemcpyAsync(intensity + nxn * off, P->d_tp_i, sizeof(double) * nxn * nb, hipMemcpyDeviceToHost,
                              P->stream));
        if ((st = harvest_status(P, nb, status ? status + off : nullptr, &worst))) return st;
    }
    if (P->profiling) harvest_timers(P);
    return worst ? (rg_status)worst : RG_OK;
}

rg_status rg_tprofile(rg_plan *P, int64_t seed0, int32_t nseed, int32_t filtered, double *tvals, double *intensity,
                      int32_t *status)
{
    return tprofile_pass(P, seed0, nullptr, nseed, filtered, tvals, intensity, status);
}

rg_status rg_tprofile_potential(rg_plan *P, const double *upot, int32_t nreal, int32_t filtered, double *tvals,
                                double *intensity, int32_t *status)
{
    if (!upot && nreal > 0) return RG_EINVAL;
    return tprofile_pass(P, 0, upot, nreal, filtered, tvals, intensity, status);
}

static rg_status wavefunction_pass(rg_plan *P, int64_t seed, const double *upot, int32_t ipos, int32_t imode,
                                   double *psi, int32_t *status)
{
    rg_status st = check_common(P, 1, RG_FP64);
    if (st) return st;
    if (!psi || imode < 1 || imode > P->nmu || ipos < 1 || ipos > P->nx) return RG_EINVAL;  // :646-654
    if (P->tp_cap < 1) return RG_EINVAL;
    // imode = 1 is the frontal mode: the middle of the plane-wave list (:657-661)
    const int imu = P->p.boundtype == RG_PERIODIC ? (imode + P->nmu / 2 - 1) % P->nmu + 1 : imode;
    DeviceGuard g(P->device);
    const size_t per = (size_t)P->nxdiso * P->ny;
    HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t), P->stream));
    const double *dup = nullptr;
    if (upot) {
        for (size_t e = 0; e < per; e++)
            if (!std::isfinite(upot[e])) return RG_ENONFINITE;
        HIPCHK(hipMemcpyAsync(P->d_upot, upot, sizeof(double) * per, hipMemcpyHostToDevice, P->stream));
        dup = P->d_upot;
    }
    const size_t need = (size_t)P->nx * P->ny;
    run_sweep_wavefunction(P, seed, dup, ipos, imu, P->d_tp_psi);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(psi, P->d_tp_psi, sizeof(double2) * need, hipMemcpyDeviceToHost, P->stream));
    int32_t worst = 0;
    if ((st = harvest_status(P, 1, status, &worst))) return st;
    if (P->profiling) harvest_timers(P);
    return worst ? (rg_status)worst : RG_OK;
}

rg_status rg_wavefunction(rg_plan *P, int64_t seed, int32_t ipos, int32_t imode, double *psi, int32_t *status)
{
    return wavefunction_pass(P, seed, nullptr, ipos, imode, psi, status);
}

rg_status rg_wavefunction_potential(rg_plan *P, const double *upot, int32_t ipos, int32_t imode, double *psi,
                                    int32_t *status)
{
    if (!upot) return RG_EINVAL;
    return wavefunction_pass(P, 0, upot, ipos, imode, psi, status);
}

rg_status rg_adjoint_reserve(rg_plan *P, int32_t nreal_max)
{
    if (!P || nreal_max < 1) return RG_EINVAL;
    if (P->np > 256) return RG_ENOTSUP;
    DeviceGuard g(P->device);
    if (P->adj_cap >= nreal_max) return RG_OK;
    for (void *b : {(void *)P->d_ck, (void *)P->d_seg, (void *)P->d_adj, (void *)P->d_loss, (void *)P->d_grad,
                    (void *)P->d_adj_work})
        if (b) (void)hipFree(b);
    P->d_ck = P->d_seg = P->d_adj = P->d_adj_work = nullptr;
    P->d_loss = P->d_grad = nullptr;
    P->adj_cap = 0;
    const int nx = P->nx;
    int seg = (int)std::ceil(std::sqrt((double)nx));
    if (seg < 1) seg = 1;
    const int nck = std::max(0, (nx - 1) / seg);  // checkpoints after slices seg, 2 seg, ... < nx
    const size_t cmat = (size_t)nreal_max * P->np * P->np;
    rg_status st;
    if ((st = dalloc(&P->d_ck, std::max<size_t>(1, 2 * (size_t)nck) * cmat)) ||
        (st = dalloc(&P->d_seg, 2 * (size_t)seg * cmat)) || (st = dalloc(&P->d_adj, 5 * cmat)) ||
        (st = dalloc(&P->d_loss, (size_t)nreal_max)) ||
        (st = dalloc(&P->d_grad, (size_t)nreal_max * P->nxdiso * P->ny)))
        return st;
    if (!rg_topvec_supported(P) && (st = dalloc(&P->d_adj_work, 2 * cmat))) return st;  // T_max above order 128
    P->adj_cap = nreal_max;
    P->adj_seg = seg;
    P->adj_nck = nck;
    return RG_OK;
}

static rg_status adjoint_check(rg_plan *P, int32_t nreal, int32_t loss_kind)
{
    rg_status st = check_common(P, nreal, RG_FP64);
    if (st) return st;
    if (loss_kind != 0 && loss_kind != 1) return RG_EINVAL;
    if (loss_kind == 1 && !rg_eig_supported(P)) return RG_ENOTSUP;  // Gram order above 256
    if (P->adj_cap < 1) return RG_EINVAL;  // rg_adjoint_reserve first
    return RG_OK;
}

rg_status rg_adjoint_potential_device(rg_plan *P, const double *upot_dev, int32_t nreal, int32_t loss_kind,
                                      double *loss_dev, double *grad_dev, int32_t *status_dev)
{
    rg_status st = adjoint_check(P, nreal, loss_kind);
    if (st) return st;
    if (nreal > 0 && (!upot_dev || !loss_dev || !grad_dev)) return RG_EINVAL;
    DeviceGuard g(P->device);
    const size_t per = (size_t)P->nxdiso * P->ny;
    const int cap = std::min(P->adj_cap, P->batch_max);
    for (int32_t off = 0; off < nreal; off += cap) {
        const int nb = std::min(cap, nreal - off);
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        run_adjoint_any(P, nb, loss_kind, upot_dev + per * off, loss_dev + off, grad_dev + per * off);
        HIPCHK(hipGetLastError());
        if (status_dev)
            HIPCHK(hipMemcpyAsync(status_dev + off, P->d_status, sizeof(int32_t) * nb, hipMemcpyDeviceToDevice, P->stream));
    }
    return RG_OK;
}

rg_status rg_adjoint_potential(rg_plan *P, const double *upot, int32_t nreal, int32_t loss_kind, double *loss,
                               double *grad, int32_t *status)
{
    rg_status st = adjoint_check(P, nreal, loss_kind);
    if (st) return st;
    if (nreal > 0 && (!upot || !loss || !grad)) return RG_EINVAL;
    DeviceGuard g(P->device);
    const size_t per = (size_t)P->nxdiso * P->ny;
    const int cap = std::min(std::min(P->adj_cap, P->upot_cap), P->batch_max);
    int32_t worst = 0;
    for (int32_t off = 0; off < nreal; off += cap) {
        const int nb = std::min(cap, nreal - off);
        for (size_t e = 0; e < per * nb; e++)
            if (!std::isfinite(upot[per * off + e])) return RG_ENONFINITE;
        HIPCHK(hipMemsetAsync(P->d_status, 0, sizeof(int32_t) * nb, P->stream));
        HIPCHK(hipMemcpyAsync(P->d_upot, upot + per * off, sizeof(double) * per * nb, hipMemcpyHostToDevice, P->stream));
        run_adjoint_any(P, nb, loss_kind, P->d_upot, P->d_loss, P->d_grad);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(loss + off, P->d_loss, sizeof(double) * nb, hipMemcpyDeviceToHost, P->stream));
        HIPCHK(hipMemcpyAsync(grad + per * off, P->d_grad, sizeof(double) * per * nb, hipMemcpyDeviceToHost, P->stream));
        if ((st = harvest_status(P, nb, status ? status + off : nullptr, &worst))) return st;
    }
    if (P->profiling) harvest_timers(P);
    return worst ? (rg_status)worst : RG_OK;
}

rg_status rg_generate_potential(rg_plan *P, int64_t seed0, int32_t nreal, double *upot)
{
    if (!P || nreal < 0 || (!upot && nreal > 0)) return RG_EINVAL;
    DeviceGuard g(P->device);
    const size_t per = (size_t)P->nxdiso * P->ny;
    for (int32_t off = 0; off < nreal; off += P->upot_cap) {
        const int nb = std::min(P->upot_cap, nreal - off);
        rg_launch_potential(P, nb, seed0 + off, P->d_upot);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(upot + per * off, P->d_upot, sizeof(double) * per * nb, hipMemcpyDeviceToHost, P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
    }
    return RG_OK;
}

rg_status rg_histogram(rg_plan *P, const double *samples, int64_t nsamp, int32_t on_device, const double *bins,
                       int32_t nbins, int64_t *counts, int64_t *missing, int64_t *dropped)
{
    if (!P || !bins || !counts || nbins < 1 || nsamp < 0 || (!samples && nsamp > 0)) return RG_EINVAL;
    if (nbins > P->hist_cap) return RG_ENOTSUP;
    for (int i = 0; i < nbins; i++)
        if (!(bins[i] <= bins[i + 1])) return RG_EINVAL;  // check_histogram, src/base_utils.f90:468-482
    DeviceGuard g(P->device);
    HIPCHK(hipMemcpyAsync(P->d_bins, bins, sizeof(double) * (nbins + 1), hipMemcpyHostToDevice, P->stream));
    HIPCHK(hipMemsetAsync(P->d_hist, 0, sizeof(unsigned long long) * (nbins + 2), P->stream));
    if (on_device) {
        if (nsamp > 0) rg_launch_histogram(P, samples, nsamp, P->d_bins, nbins, P->d_hist);
    } else {
        // host samples are staged through the explicit-potential buffer in chunks
        const int64_t chunk = (int64_t)P->upot_cap * P->nxdiso * P->ny;
        for (int64_t off = 0; off < nsamp; off += chunk) {
            const int64_t nc = std::min(chunk, nsamp - off);
            HIPCHK(hipMemcpyAsync(P->d_upot, samples + off, sizeof(double) * nc, hipMemcpyHostToDevice, P->stream));
            rg_launch_histogram(P, P->d_upot, nc, P->d_bins, nbins, P->d_hist);
        }
    }
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> h(nbins + 2);
    HIPCHK(hipMemcpyAsync(h.data(), P->d_hist, sizeof(unsigned long long) * (nbins + 2), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    for (int i = 0; i < nbins; i++) counts[i] = (int64_t)h[i];
    if (missing) *missing = (int64_t)h[nbins];
    if (dropped) *dropped = (int64_t)h[nbins + 1];
    if (P->profiling) harvest_timers(P);
    return RG_OK;
}

rg_status rg_plan_stream(rg_plan *P, void **stream)
{
    if (!P || !stream) return RG_EINVAL;
    *stream = (void *)P->stream;
    return RG_OK;
}

rg_status rg_plan_synchronize(rg_plan *P)
{
    if (!P) return RG_EINVAL;
    DeviceGuard g(P->device);
    HIPCHK(hipStreamSynchronize(P->stream));
    HIPCHK(hipGetLastError());
    if (P->profiling) harvest_timers(P);
    return RG_OK;
}

rg_status rg_plan_profile(rg_plan *P, int32_t enable)
{
    if (!P) return RG_EINVAL;
    DeviceGuard g(P->device);
    (void)hipStreamSynchronize(P->stream);
    harvest_timers(P);
    for (auto &t : P->timers) { t.ms = 0.0; t.launches = 0; }
    P->profiling = enable != 0;
    return RG_OK;
}

rg_status rg_plan_set_option(rg_plan *P, const char *key, int32_t value)
{
    if (!P || !key) return RG_EINVAL;
    DeviceGuard g(P->device);
    (void)hipStreamSynchronize(P->stream);
    if (std::strcmp(key, "inverse_fp64") == 0) {
        if (value < RG_INV_REGISTER || value > RG_INV_SYM) return RG_EINVAL;
        P->inv_fp64 = value;
    } else if (std::strcmp(key, "inverse_mixed") == 0) {
        if (value < RG_INV_LOOKAHEAD || value > RG_INV_SYM) return RG_EINVAL;
        P->inv_mixed = value;
    } else if (std::strcmp(key, "dyson") == 0) {
        if (value < 0 || value > 3) return RG_EINVAL;
        P->gemm_wg = value;
    } else if (std::strcmp(key, "dyson_fp64") == 0) {
        if (value < 0 || value > 2) return RG_EINVAL;
        P->dyson64 = value;
    } else if (std::strcmp(key, "project_parts") == 0) {
        if (value < 1 || value > kMaxCqParts) return RG_EINVAL;
        P->cq_parts = value;
    } else if (std::strcmp(key, "project_wide_min") == 0) {
        if (value < 1) return RG_EINVAL;
        P->project_wide_min = value;
    } else if (std::strcmp(key, "refine") == 0) {
        if (value < 0 || value > 8) return RG_EINVAL;
        P->refine = value;
    } else if (std::strcmp(key, "refine_cap") == 0) {
        if (value < 0) return RG_EINVAL;
        P->refine_cap = value;
        rg_refine_release(P);  // the next refined call sizes its buffers again
    } else if (std::strcmp(key, "lanes") == 0) {
        if (value != 1 && value != 2) return RG_EINVAL;
        if (value == 2) {
            rg_status st = ensure_lanes(P);
            if (st) return st;
        }
        P->lanes = value;
    } else if (std::strcmp(key, "eigen") == 0) {
        if (value < RG_EIG_GRAM || value > RG_EIG_AUTO) return RG_EINVAL;
        P->eigen = value;
    } else if (std::strcmp(key, "profile_mask") == 0) {
        P->profile_mask = value;
    } else if (std::strcmp(key, "graph") == 0) {
        if (value != 0 && value != 1) return RG_EINVAL;
        P->use_graph = value == 1;
    } else if (std::strcmp(key, "adjoint_precision") == 0) {
        if (value != RG_FP64 && value != RG_MIXED_FP32) return RG_EINVAL;
        P->adj_prec = value;
    } else {
        return RG_EINVAL;
    }
    return RG_OK;
}

rg_status rg_plan_get_option(const rg_plan *P, const char *key, int32_t *value)
{
    if (!P || !key || !value) return RG_EINVAL;
    if (std::strcmp(key, "inverse_fp64") == 0) *value = P->inv_fp64;
    else if (std::strcmp(key, "inverse_mixed") == 0) *value = P->inv_mixed;
    else if (std::strcmp(key, "dyson_fp64") == 0) *value = P->dyson64;
    else if (std::strcmp(key, "dyson") == 0) *value = P->gemm_wg;
    else if (std::strcmp(key, "project_parts") == 0) *value = P->cq_parts;
    else if (std::strcmp(key, "project_wide_min") == 0) *value = P->project_wide_min;
    else if (std::strcmp(key, "refine") == 0) *value = P->refine;
    else if (std::strcmp(key, "refine_cap") == 0) *value = P->refine_cap;
    else if (std::strcmp(key, "lanes") == 0) *value = P->lanes;
    else if (std::strcmp(key, "eigen") == 0) *value = P->eigen;
    else if (std::strcmp(key, "profile_mask") == 0) *value = P->profile_mask;
    else if (std::strcmp(key, "graph") == 0) *value = P->use_graph ? 1 : 0;
    else if (std::strcmp(key, "adjoint_precision") == 0) *value = P->adj_prec;
    else return RG_EINVAL;
    return RG_OK;
}

rg_status rg_plan_kernel_times(rg_plan *P, int32_t cap, char *names, double *ms, int64_t *launches, int32_t *count)
{
    if (!P || !count) return RG_EINVAL;
    DeviceGuard g(P->device);
    (void)hipStreamSynchronize(P->stream);
    harvest_timers(P);
    int k = 0;
    for (int i = 0; i < RGK_COUNT && k < cap; i++, k++) {
        if (names) {
            std::memset(names + 32 * k, 0, 32);
            std::strncpy(names + 32 * k, P->timers[i].name.c_str(), 31);
        }
        if (ms) ms[k] = P->timers[i].ms;
        if (launches) launches[k] = P->timers[i].launches;
    }
    *count = k;
    return RG_OK;
}

const char *rg_status_string(int32_t code)
{
    switch (code) {
        case RG_OK: return "RecurGreen: success";
        case RG_EINVAL: return "RecurGreen: Invalid arguments, aborting...";
        case RG_ESINGULAR: return "RecurGreen: zgetrf: Matrix is numerically singular, aborting...";
        case RG_ENOCONV: return "RecurGreen: zgesvd: Failed to find the singular values, aborting...";
        case RG_ENOMEM: return "RecurGreen: device or host allocation failed";
        case RG_EDEVICE: return "RecurGreen: HIP runtime error";
        case RG_ENOTSUP: return "RecurGreen: combination not supported by this build";
        case RG_ENONFINITE: return "RecurGreen: non-finite value in the potential or in G(L;0)";
        default: return "RecurGreen: unknown status";
    }
}

const char *rg_version(void) { return "recurgreen-hip 0.1 gfx950"; }

}  // extern "C"
