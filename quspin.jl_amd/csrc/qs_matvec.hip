// qs_matvec.hip -- matrix-free sector Hamiltonian action y = beta*y + alpha*M x.
//
// Reference: QuantumLinearOperator mul! (operators.jl:3147-3179) with
// _apply_spin_terms! (operators.jl:2591-2636): lift x to the parent
// particle sector with the explicit projector, apply every coupling with a
// hash lookup per (state, coupling), project back.  That is M = P' H P.
//
// Here M is applied in pull form on the sector itself.  With the column
// |c> = sum_g conj(chi(g)) g|s_c> / sqrt(D_c) (s_c the orbit minimum; the
// reference's columns are these up to rounding) and H replaced by its group
// average (P'HP = P'H_avg P), one output row c' with canonical state s is
//   y[c'] = sum_u <s|H|u> chi(h_u) sqrt(D_c'/D_c) x[c(u)],   h_u u = s_c(u),
// where u runs over the states the adjoint couplings reach from s.  One lane
// owns one row: no atomics, a fixed summation order, so results do not
// depend on what else runs on the device.
#include <algorithm>

#include "qs_device.h"

namespace qs {

namespace {

constexpr int kMvBlock = 256;
constexpr int kMaxSamples = 4096;

struct MVArgs {
    int64_t N;
    int L, nup, zchar;
    uint64_t mask;
    GroupDev g;
    const void *reps;
    const double *inv_nrm;
    const void *samp;
    int64_t nsample;
    int shift;
    const uint64_t *binom;
    const TermDev *diag;
    int ndiag;
    const TermDev *off;
    const GroupOfTerms *groups;
    int ngroups;
    const double *user_diag;   // interleaved complex or null
    int user_diag_conj;
    double are, aim, bre, bim;
    int conj_x, conj_y;
    int beta_zero;
    double *dot_partial;   // optional: per-block FP64 partials of <x, y>
};

// Amplitude of one compiled term on state s: A * (-1)^popcount(s & smask),
// or zero when the conditions fail.
template <typename S, bool CC>
__device__ __forceinline__ void term_amp(const TermDev &t, S s, double &re, double &im) {
    if ((s & (S)t.cmask) != (S)t.cval) return;
    double sg = (popc(s & (S)t.smask) & 1) ? -1.0 : 1.0;
    re += sg * t.are;
    if (CC) im += sg * t.aim;
}

template <typename S>
__device__ __forceinline__ int64_t search_rep(const S *reps, const S *lsamp, int64_t nsample, int shift,
                                              int64_t N, S r) {
    if (r < lsamp[0]) return -1;
    int64_t lo = 0, hi = nsample;
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (lsamp[mid] <= r) lo = mid; else hi = mid;
    }
    int64_t a = lo << shift, b = (lo + 1) << shift;
    if (b > N) b = N;
    while (b - a > 1) {
        int64_t mid = (a + b) >> 1;
        if (reps[mid] <= r) a = mid; else b = mid;
    }
    return reps[a] == r ? a : -1;
}

// MODE: 0 full, 1 z-only, 2 stored; TRACK: sector characters not all one;
// CC: complex coefficients / characters.
template <typename S, typename V, bool CC, int MODE, bool TRACK>
__global__ __launch_bounds__(kMvBlock) void k_matvec(MVArgs A, const V *__restrict__ x, V *__restrict__ y) {
    using T = VTraits<V>;
    __shared__ double s_red[2 * kMvBlock];
    double dot_r = 0, dot_i = 0;
    __shared__ S s_samp[MODE == MODE_STORED ? kMaxSamples : 1];
    __shared__ uint32_t s_binom[MODE == MODE_STORED ? 1 : 33 * 65];
    if (MODE == MODE_STORED) {
        const S *samp = (const S *)A.samp;
        for (int i = threadIdx.x; i < A.nsample; i += kMvBlock) s_samp[i] = samp[i];
    } else if (sizeof(S) == 4 && A.nup >= 0) {
        for (int i = threadIdx.x; i < 33 * 65; i += kMvBlock) s_binom[i] = (uint32_t)A.binom[i];
    }
    __syncthreads();
    const S mask = (S)A.mask;
    const S *reps = (const S *)A.reps;
    for (int64_t row = (int64_t)blockIdx.x * kMvBlock + threadIdx.x; row < A.N;
         row += (int64_t)gridDim.x * kMvBlock) {
        S s;
        if (MODE == MODE_STORED) {
            s = reps[row];
        } else if (A.nup >= 0) {
            if (sizeof(S) == 4) s = comb_unrank<S, uint32_t>(row, A.L, A.nup, s_binom);
            else s = comb_unrank<S, uint64_t>(row, A.L, A.nup, A.binom);
        } else {
            s = (S)row;
        }
        double own_scale = 1.0;
        if (MODE == MODE_STORED) own_scale = 1.0 / A.inv_nrm[row];
        double xr, xi;
        T::load(x, row, xr, xi);
        if (A.conj_x) xi = -xi;
        // diagonal couplings
        double dr = 0, di = 0;
        for (int k = 0; k < A.ndiag; k++) term_amp<S, CC>(A.diag[k], s, dr, di);
        di = -di;   // conj(amp): the terms are the adjoint couplings
        if (A.user_diag) {
            dr += A.user_diag[2 * row];
            di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
        }
        double ar = dr * xr - di * xi;
        double ai = dr * xi + di * xr;
        // off-diagonal couplings, grouped by flip mask
        for (int gi = 0; gi < A.ngroups; gi++) {
            const GroupOfTerms gr = A.groups[gi];
            double mr = 0, mi = 0;
            for (int k = gr.t0; k < gr.t1; k++) term_amp<S, CC>(A.off[k], s, mr, mi);
            if (mr == 0.0 && mi == 0.0) continue;
            mi = -mi;
            S u = s ^ (S)gr.flip;
            int64_t col;
            double cr = 1.0, ci = 0.0;
            if (MODE == MODE_FULL) {
                if (A.nup >= 0) {
                    if (popc(u) != A.nup) continue;
                    col = (sizeof(S) == 4) ? comb_rank<S, uint32_t>(u, s_binom) : comb_rank<S, uint64_t>(u, A.binom);
                } else {
                    col = (int64_t)u;
                }
            } else if (MODE == MODE_ZONLY) {
                if (A.nup >= 0 && popc(u) != A.nup) continue;
                S r = u;
                if ((u >> (A.L - 1)) & 1) {
                    r = ~u & mask;
                    cr = (double)A.zchar;
                }
                if (A.nup >= 0)
                    col = (sizeof(S) == 4) ? comb_rank<S, uint32_t>(r, s_binom) : comb_rank<S, uint64_t>(r, A.binom);
                else
                    col = (int64_t)r;
            } else {
                int bt, bp;
                S r = canonical<S, TRACK>(A.g, u, mask, bt, bp);
                col = search_rep<S>(reps, s_samp, A.nsample, A.shift, A.N, r);
                if (col < 0) continue;
                if (TRACK) {
                    double lr = A.g.lam_re[bt], li = A.g.lam_im[bt];
                    double pr = A.g.pch_re[bp], pi = A.g.pch_im[bp];
                    cr = lr * pr - li * pi;
                    ci = lr * pi + li * pr;
                }
                double f = A.inv_nrm[col] * own_scale;
                cr *= f;
                ci *= f;
            }
            double wr = mr * cr - mi * ci;
            double wi = mr * ci + mi * cr;
            double vr, vi;
            T::load(x, col, vr, vi);
            if (A.conj_x) vi = -vi;
            ar += wr * vr - wi * vi;
            ai += wr * vi + wi * vr;
        }
        // y = beta*y + alpha*acc (with optional conjugation of y on entry / exit)
        double orr = A.are * ar - A.aim * ai;
        double oi = A.are * ai + A.aim * ar;
        if (!A.beta_zero) {
            double yr, yi;
            T::load(y, row, yr, yi);
            if (A.conj_y) yi = -yi;
            orr += A.bre * yr - A.bim * yi;
            oi += A.bre * yi + A.bim * yr;
        }
        if (A.conj_y) oi = -oi;
        if (A.dot_partial) {
            // FP64 <x, y_row> of the unrounded result (Rayleigh quotients of FP32-stored vectors)
            dot_r += xr * orr + xi * oi;
            dot_i += xr * oi - xi * orr;
        }
        T::store(y, row, orr, oi);
    }
    if (A.dot_partial) {
        s_red[threadIdx.x] = dot_r;
        s_red[kMvBlock + threadIdx.x] = dot_i;
        __syncthreads();
        for (int st = kMvBlock / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st) {
                s_red[threadIdx.x] += s_red[threadIdx.x + st];
                s_red[kMvBlock + threadIdx.x] += s_red[kMvBlock + threadIdx.x + st];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            A.dot_partial[2 * blockIdx.x] = s_red[0];
            A.dot_partial[2 * blockIdx.x + 1] = s_red[kMvBlock];
        }
    }
}

template <typename S, typename V, bool CC>
void dispatch_mode(const MVArgs &a, int mode, bool track, const void *x, void *y, int grid, hipStream_t s) {
    const V *xv = (const V *)x;
    V *yv = (V *)y;
    if (mode == MODE_FULL)
        hipLaunchKernelGGL((k_matvec<S, V, CC, MODE_FULL, false>), dim3(grid), dim3(kMvBlock), 0, s, a, xv, yv);
    else if (mode == MODE_ZONLY)
        hipLaunchKernelGGL((k_matvec<S, V, CC, MODE_ZONLY, false>), dim3(grid), dim3(kMvBlock), 0, s, a, xv, yv);
    else if (track)
        hipLaunchKernelGGL((k_matvec<S, V, CC, MODE_STORED, true>), dim3(grid), dim3(kMvBlock), 0, s, a, xv, yv);
    else
        hipLaunchKernelGGL((k_matvec<S, V, CC, MODE_STORED, false>), dim3(grid), dim3(kMvBlock), 0, s, a, xv, yv);
}

template <typename S, typename V>
void dispatch_cc(const MVArgs &a, bool cc, int mode, bool track, const void *x, void *y, int grid, hipStream_t s) {
    if constexpr (VTraits<V>::cplx) {
        if (cc) dispatch_mode<S, V, true>(a, mode, track, x, y, grid, s);
        else dispatch_mode<S, V, false>(a, mode, track, x, y, grid, s);
    } else {
        dispatch_mode<S, V, false>(a, mode, track, x, y, grid, s);
    }
}

template <typename S>
void dispatch_v(const MVArgs &a, int dtype, bool cc, int mode, bool track, const void *x, void *y, int grid,
                hipStream_t s) {
    switch (dtype) {
    case QS_F64: dispatch_cc<S, double>(a, cc, mode, track, x, y, grid, s); break;
    case QS_C128: dispatch_cc<S, cplx_d>(a, cc, mode, track, x, y, grid, s); break;
    case QS_F32: dispatch_cc<S, float>(a, cc, mode, track, x, y, grid, s); break;
    case QS_C64: dispatch_cc<S, cplx_f>(a, cc, mode, track, x, y, grid, s); break;
    default: fail(QS_EARG, "unknown dtype");
    }
}

GroupDev group_dev(const GroupHost &h) {
    GroupDev d;
    memset(&d, 0, sizeof d);
    d.L = h.L;
    d.a = h.a;
    d.nT = h.nT;
    d.npoint = h.npoint;
    for (int p = 0; p < h.npoint; p++) {
        d.prev[p] = h.prev[p];
        d.pflip[p] = h.pflip[p];
        d.pch_re[p] = h.pchar[p].real();
        d.pch_im[p] = h.pchar[p].imag();
    }
    for (int t = 0; t < h.nT; t++) {
        d.lam_re[t] = h.lam[t].real();
        d.lam_im[t] = h.lam[t].imag();
    }
    return d;
}

}  // namespace

int matvec_grid(int64_t n) {
    int64_t blocks = (n + kMvBlock - 1) / kMvBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, 65536));
}

void launch_matvec(const OpHost &op, int dtype, const void *x, void *y, double are, double aim, double bre,
                   double bim, bool adjoint, bool conj_x, bool conj_y, hipStream_t s, double *dot_partial) {
    const BasisHost &b = *op.basis;
    MVArgs a;
    memset(&a, 0, sizeof a);
    a.N = b.Ns;
    a.L = b.L;
    a.nup = b.nup;
    a.zchar = b.zchar;
    a.mask = full_mask(b.L);
    a.g = group_dev(b.group);
    a.reps = b.use64 ? (const void *)b.reps64.get() : (const void *)b.reps32.get();
    a.inv_nrm = b.inv_nrm.get();
    a.samp = b.use64 ? (const void *)b.samp64.get() : (const void *)b.samp32.get();
    a.nsample = b.nsample;
    a.shift = b.sample_shift;
    a.binom = b.binom.get();
    // plain products use the adjoint couplings (pull form), adjoint products the couplings
    if (!adjoint) {
        a.diag = op.d_diag.get();
        a.ndiag = (int)op.diag_terms.size();
        a.off = op.d_off.get();
        a.groups = op.d_groups.get();
        a.ngroups = (int)op.groups.size();
    } else {
        a.diag = op.d_diag_adj.get();
        a.ndiag = (int)op.diag_terms_adj.size();
        a.off = op.d_off_adj.get();
        a.groups = op.d_groups_adj.get();
        a.ngroups = (int)op.groups_adj.size();
    }
    a.user_diag = op.has_user_diag ? op.d_user_diag.get() : nullptr;
    a.user_diag_conj = adjoint ? 1 : 0;
    a.are = are;
    a.aim = aim;
    a.bre = bre;
    a.bim = bim;
    a.beta_zero = (bre == 0.0 && bim == 0.0);
    a.conj_x = conj_x ? 1 : 0;
    a.conj_y = conj_y ? 1 : 0;
    a.dot_partial = dot_partial;
    bool track = false;
    for (int p = 0; p < b.group.npoint; p++)
        if (b.group.pchar[p] != cd(1, 0)) track = true;
    for (int t = 0; t < b.group.nT; t++)
        if (b.group.lam[t] != cd(1, 0)) track = true;
    bool cc = op.is_complex != 0;
    if (cc && !dtype_complex(dtype)) fail(QS_EARG, "a complex operator needs a complex vector dtype");
    int grid = matvec_grid(b.Ns);
    if (b.use64) dispatch_v<uint64_t>(a, dtype, cc, b.mode, track, x, y, grid, s);
    else dispatch_v<uint32_t>(a, dtype, cc, b.mode, track, x, y, grid, s);
    QS_HIP(hipGetLastError());
}

}  // namespace qs
