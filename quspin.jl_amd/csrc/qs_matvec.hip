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
// where u runs over the states the adjoint couplings reach from s.
//
// Work decomposition: LPR lanes share one row (LPR in {1, 2, 4, 8}, picked
// per operator); each lane evaluates the amplitude of its share of the
// flip groups, keeps the live ones in a bit mask and walks only those, so
// lanes do not idle on dead bonds.  The LPR partial sums are combined with a
// fixed xor-shuffle tree and one lane writes y: no atomics and a fixed
// summation order, so a result never depends on what else runs on the device.
// Column lookup: combinadic rank (no symmetry / spin inversion only),
// updated incrementally for two-site hops; otherwise the orbit minimum (32-bit
// funnel-shift windows) located in the sorted canonical states through a
// binary search whose top levels sit in LDS.  Blocks are renumbered so each
// XCD owns a contiguous row range and keeps its own L2 warm.
#include <algorithm>

#include "qs_device.h"

namespace qs {

namespace {

constexpr int kMvBlock = 256;
constexpr int kMaxLdsGroups = 256;
constexpr int kMaxLdsTerms = 512;

struct MVArgs {
    int64_t N;
    int L, nup, zchar;
    uint64_t mask;
    GroupDev g;
    const void *reps;
    const double *inv_nrm;
    const void *samp;
    int nsample;
    int shift;
    const uint64_t *binom;
    const TermDev *diag;
    int ndiag;
    const TermDev *off;
    int noff;
    const GroupOfTerms *groups;
    int ngroups;
    const double *user_diag;   // interleaved complex or null
    int user_diag_conj;
    double are, aim, bre, bim;
    int conj_x, conj_y;
    int beta_zero;
    double *dot_partial;   // optional: per-block FP64 partials of <x, y>
    // dynamic LDS layout (byte offsets)
    int off_samp, off_binom, off_grp, off_term, off_red;
    int lds_groups, lds_terms;
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
__device__ __forceinline__ int64_t search_rep(const S *reps, const S *lsamp, int nsample, int shift, int64_t N, S r) {
    if (r < lsamp[0]) return -1;
    int lo = 0, hi = nsample;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (lsamp[mid] <= r) lo = mid; else hi = mid;
    }
    int64_t a = (int64_t)lo << shift, b = (int64_t)(lo + 1) << shift;
    if (b > N) b = N;
    while (b - a > 1) {
        int64_t mid = (a + b) >> 1;
        if (reps[mid] <= r) a = mid; else b = mid;
    }
    return reps[a] == r ? a : -1;
}

// Colex rank change when the one particle on bits {i, j} (i < j) hops to
// the other bit: the moved particle changes index, the m particles strictly
// between them shift by one.
template <typename S, typename B>
__device__ __forceinline__ int64_t rank_delta(S s, int i, int j, const B *binom) {
    const int c = popc(s & (((S)1 << i) - 1));
    const S between = s & ((((S)1 << j) - 1) ^ ((((S)1 << i) << 1) - 1));
    const int m = popc(between);
    int64_t d = (int64_t)binom[j * 65 + c + m + 1] - (int64_t)binom[i * 65 + c + 1];
    S b = between;
    int l = 1;
    while (b) {
        int q = ctz(b);
        d += (int64_t)binom[q * 65 + c + l] - (int64_t)binom[q * 65 + c + l + 1];
        l++;
        b &= b - 1;
    }
    return ((s >> i) & 1) ? d : -d;
}

// hardware places block b on XCD b % 8; give each XCD a contiguous range
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int xcd = b & 7, idx = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

template <typename S, typename V, bool CC, int MODE, bool TRACK, int LPR>
__global__ __launch_bounds__(kMvBlock) void k_matvec(MVArgs A, const V *__restrict__ x, V *__restrict__ y) {
    using T = VTraits<V>;
    extern __shared__ __align__(16) char smem[];
    S *s_samp = (S *)(smem + A.off_samp);
    uint32_t *s_binom = (uint32_t *)(smem + A.off_binom);
    GroupOfTerms *s_grp = (GroupOfTerms *)(smem + A.off_grp);
    TermDev *s_term = (TermDev *)(smem + A.off_term);
    double *s_red = (double *)(smem + A.off_red);
    constexpr bool kLdsBinom = (MODE != MODE_STORED) && sizeof(S) == 4;
    if (MODE == MODE_STORED) {
        const S *samp = (const S *)A.samp;
        for (int i = threadIdx.x; i < A.nsample; i += kMvBlock) s_samp[i] = samp[i];
    }
    if (kLdsBinom && A.nup >= 0)
        for (int i = threadIdx.x; i < 33 * 65; i += kMvBlock) s_binom[i] = (uint32_t)A.binom[i];
    if (A.lds_groups)
        for (int i = threadIdx.x; i < A.ngroups; i += kMvBlock) s_grp[i] = A.groups[i];
    if (A.lds_terms)
        for (int i = threadIdx.x; i < A.noff; i += kMvBlock) s_term[i] = A.off[i];
    __syncthreads();
    const GroupOfTerms *G = A.lds_groups ? s_grp : A.groups;
    const TermDev *TT = A.lds_terms ? s_term : A.off;
    const S mask = (S)A.mask;
    const S *reps = (const S *)A.reps;
    const int sub = threadIdx.x % LPR;
    constexpr int kRows = kMvBlock / LPR;
    const int vb = xcd_block(blockIdx.x, gridDim.x);
    double dot_r = 0, dot_i = 0;
    for (int64_t base = (int64_t)vb * kRows; base < A.N; base += (int64_t)gridDim.x * kRows) {
        const int64_t row = base + threadIdx.x / LPR;
        const bool valid = row < A.N;
        double ar = 0, ai = 0, xr = 0, xi = 0;
        if (valid) {
            S s;
            if (MODE == MODE_STORED) {
                s = reps[row];
            } else if (A.nup >= 0) {
                if (kLdsBinom) s = comb_unrank<S, uint32_t>(row, A.L, A.nup, s_binom);
                else s = comb_unrank<S, uint64_t>(row, A.L, A.nup, A.binom);
            } else {
                s = (S)row;
            }
            double own_scale = 1.0;
            if (MODE == MODE_STORED) own_scale = 1.0 / A.inv_nrm[row];
            if (sub == 0) {
                T::load(x, row, xr, xi);
                if (A.conj_x) xi = -xi;
                double dr = 0, di = 0;
                for (int k = 0; k < A.ndiag; k++) term_amp<S, CC>(A.diag[k], s, dr, di);
                di = -di;   // conj(amp): the terms are the adjoint couplings
                if (A.user_diag) {
                    dr += A.user_diag[2 * row];
                    di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
                }
                ar = dr * xr - di * xi;
                ai = dr * xi + di * xr;
            }
            for (int g0 = sub; g0 < A.ngroups; g0 += 64 * LPR) {
                // live flip groups of this lane's share
                uint64_t live = 0;
                for (int jj = 0; jj < 64; jj++) {
                    const int gi = g0 + jj * LPR;
                    if (gi >= A.ngroups) break;
                    const GroupOfTerms gr = G[gi];
                    double mr = 0, mi = 0;
                    for (int k = gr.t0; k < gr.t1; k++) term_amp<S, CC>(TT[k], s, mr, mi);
                    if (mr != 0.0 || mi != 0.0) live |= 1ULL << jj;
                }
                while (live) {
                    const int jj = __builtin_ctzll(live);
                    live &= live - 1;
                    const GroupOfTerms gr = G[g0 + jj * LPR];
                    double mr = 0, mi = 0;
                    for (int k = gr.t0; k < gr.t1; k++) term_amp<S, CC>(TT[k], s, mr, mi);
                    mi = -mi;
                    const S u = s ^ (S)gr.flip;
                    int64_t col;
                    double cr = 1.0, ci = 0.0;
                    if (MODE == MODE_FULL) {
                        if (A.nup >= 0) {
                            if (popc(u) != A.nup) continue;
                            if (gr.lo >= 0) {
                                col = row + (kLdsBinom ? rank_delta<S, uint32_t>(s, gr.lo, gr.hi, s_binom)
                                                       : rank_delta<S, uint64_t>(s, gr.lo, gr.hi, A.binom));
                            } else {
                                col = kLdsBinom ? comb_rank<S, uint32_t>(u, s_binom) : comb_rank<S, uint64_t>(u, A.binom);
                            }
                        } else {
                            col = (int64_t)u;
                        }
                    } else if (MODE == MODE_ZONLY) {
                        if (A.nup >= 0 && popc(u) != A.nup) continue;
                        if ((u >> (A.L - 1)) & 1) {
                            const S r = ~u & mask;
                            cr = (double)A.zchar;
                            if (A.nup >= 0)
                                col = kLdsBinom ? comb_rank<S, uint32_t>(r, s_binom) : comb_rank<S, uint64_t>(r, A.binom);
                            else
                                col = (int64_t)r;
                        } else if (A.nup >= 0) {
                            if (gr.lo >= 0)
                                col = row + (kLdsBinom ? rank_delta<S, uint32_t>(s, gr.lo, gr.hi, s_binom)
                                                       : rank_delta<S, uint64_t>(s, gr.lo, gr.hi, A.binom));
                            else
                                col = kLdsBinom ? comb_rank<S, uint32_t>(u, s_binom) : comb_rank<S, uint64_t>(u, A.binom);
                        } else {
                            col = (int64_t)u;
                        }
                    } else {
                        int bt, bp;
                        S r;
                        if constexpr (sizeof(S) == 4) r = canonical32<TRACK>(A.g, u, bt, bp);
                        else r = canonical<S, TRACK>(A.g, u, mask, bt, bp);
                        col = search_rep<S>(reps, s_samp, A.nsample, A.shift, A.N, r);
                        if (col < 0) continue;
                        if (TRACK) {
                            const double lr = A.g.lam_re[bt], li = A.g.lam_im[bt];
                            const double pr = A.g.pch_re[bp], pi = A.g.pch_im[bp];
                            cr = lr * pr - li * pi;
                            ci = lr * pi + li * pr;
                        }
                        const double f = A.inv_nrm[col] * own_scale;
                        cr *= f;
                        ci *= f;
                    }
                    const double wr = mr * cr - mi * ci;
                    const double wi = mr * ci + mi * cr;
                    double vr, vi;
                    T::load(x, col, vr, vi);
                    if (A.conj_x) vi = -vi;
                    ar += wr * vr - wi * vi;
                    ai += wr * vi + wi * vr;
                }
            }
        }
        // fixed-order combination of the LPR partial sums
#pragma unroll
        for (int o = LPR / 2; o >= 1; o >>= 1) {
            ar += __shfl_xor(ar, o);
            ai += __shfl_xor(ai, o);
        }
        if (valid && sub == 0) {
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

struct Launch {
    int grid;
    size_t lds;
    hipStream_t s;
};

template <typename S, typename V, bool CC, int MODE, bool TRACK>
void dispatch_lpr(const MVArgs &a, int lpr, const void *x, void *y, const Launch &L) {
    const V *xv = (const V *)x;
    V *yv = (V *)y;
    switch (lpr) {
    case 1: hipLaunchKernelGGL((k_matvec<S, V, CC, MODE, TRACK, 1>), dim3(L.grid), dim3(kMvBlock), L.lds, L.s, a, xv, yv); break;
    case 2: hipLaunchKernelGGL((k_matvec<S, V, CC, MODE, TRACK, 2>), dim3(L.grid), dim3(kMvBlock), L.lds, L.s, a, xv, yv); break;
    case 4: hipLaunchKernelGGL((k_matvec<S, V, CC, MODE, TRACK, 4>), dim3(L.grid), dim3(kMvBlock), L.lds, L.s, a, xv, yv); break;
    default: hipLaunchKernelGGL((k_matvec<S, V, CC, MODE, TRACK, 8>), dim3(L.grid), dim3(kMvBlock), L.lds, L.s, a, xv, yv); break;
    }
}

template <typename S, typename V, bool CC>
void dispatch_mode(const MVArgs &a, int mode, bool track, int lpr, const void *x, void *y, const Launch &L) {
    if (mode == MODE_FULL) dispatch_lpr<S, V, CC, MODE_FULL, false>(a, lpr, x, y, L);
    else if (mode == MODE_ZONLY) dispatch_lpr<S, V, CC, MODE_ZONLY, false>(a, lpr, x, y, L);
    else if (track) dispatch_lpr<S, V, CC, MODE_STORED, true>(a, lpr, x, y, L);
    else dispatch_lpr<S, V, CC, MODE_STORED, false>(a, lpr, x, y, L);
}

template <typename S, typename V>
void dispatch_cc(const MVArgs &a, bool cc, int mode, bool track, int lpr, const void *x, void *y, const Launch &L) {
    if constexpr (VTraits<V>::cplx) {
        if (cc) dispatch_mode<S, V, true>(a, mode, track, lpr, x, y, L);
        else dispatch_mode<S, V, false>(a, mode, track, lpr, x, y, L);
    } else {
        dispatch_mode<S, V, false>(a, mode, track, lpr, x, y, L);
    }
}

template <typename S>
void dispatch_v(const MVArgs &a, int dtype, bool cc, int mode, bool track, int lpr, const void *x, void *y,
                const Launch &L) {
    switch (dtype) {
    case QS_F64: dispatch_cc<S, double>(a, cc, mode, track, lpr, x, y, L); break;
    case QS_C128: dispatch_cc<S, cplx_d>(a, cc, mode, track, lpr, x, y, L); break;
    case QS_F32: dispatch_cc<S, float>(a, cc, mode, track, lpr, x, y, L); break;
    case QS_C64: dispatch_cc<S, cplx_f>(a, cc, mode, track, lpr, x, y, L); break;
    default: fail(QS_EARG, "unknown dtype");
    }
}

}  // namespace

int default_lanes_per_row(const BasisHost &b) { return b.mode == MODE_STORED ? 4 : 1; }

int matvec_grid_lpr(int64_t n, int lpr) {
    const int rows = kMvBlock / lpr;
    int64_t blocks = (n + rows - 1) / rows;
    return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, (int64_t)1 << 22));
}

int matvec_grid(int64_t n) { return matvec_grid_lpr(n, 8); }

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
    a.g = group_to_dev(b.group);
    a.reps = b.use64 ? (const void *)b.reps64.get() : (const void *)b.reps32.get();
    a.inv_nrm = b.inv_nrm.get();
    a.samp = b.use64 ? (const void *)b.samp64.get() : (const void *)b.samp32.get();
    a.nsample = (int)b.nsample;
    a.shift = b.sample_shift;
    a.binom = b.binom.get();
    // plain products use the adjoint couplings (pull form), adjoint products the couplings
    const std::vector<CTerm> &dg = adjoint ? op.diag_terms_adj : op.diag_terms;
    const std::vector<CTerm> &of = adjoint ? op.off_terms_adj : op.off_terms;
    const std::vector<GroupOfTerms> &gr = adjoint ? op.groups_adj : op.groups;
    a.diag = adjoint ? op.d_diag_adj.get() : op.d_diag.get();
    a.ndiag = (int)dg.size();
    a.off = adjoint ? op.d_off_adj.get() : op.d_off.get();
    a.noff = (int)of.size();
    a.groups = adjoint ? op.d_groups_adj.get() : op.d_groups.get();
    a.ngroups = (int)gr.size();
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
    const bool cc = op.is_complex != 0;
    if (cc && !dtype_complex(dtype)) fail(QS_EARG, "a complex operator needs a complex vector dtype");
    int lpr = op.lanes_per_row > 0 ? op.lanes_per_row : default_lanes_per_row(b);
    if (dot_partial) lpr = 8;   // the partial-sum grid is sized by matvec_grid
    // dynamic LDS layout
    size_t off = 0;
    auto place = [&](size_t bytes) {
        size_t o = (off + 15) & ~(size_t)15;
        off = o + bytes;
        return (int)o;
    };
    const size_t ssz = b.use64 ? 8 : 4;
    a.off_samp = place(b.mode == MODE_STORED ? ssz * (size_t)b.nsample : 0);
    a.off_binom = place((b.mode != MODE_STORED && !b.use64 && b.nup >= 0) ? sizeof(uint32_t) * 33 * 65 : 0);
    a.lds_groups = a.ngroups <= kMaxLdsGroups;
    a.lds_terms = a.noff <= kMaxLdsTerms;
    a.off_grp = place(a.lds_groups ? sizeof(GroupOfTerms) * a.ngroups : 0);
    a.off_term = place(a.lds_terms ? sizeof(TermDev) * a.noff : 0);
    a.off_red = place(dot_partial ? sizeof(double) * 2 * kMvBlock : 0);
    Launch L{matvec_grid_lpr(b.Ns, lpr), off, s};
    if (b.use64) dispatch_v<uint64_t>(a, dtype, cc, b.mode, track, lpr, x, y, L);
    else dispatch_v<uint32_t>(a, dtype, cc, b.mode, track, lpr, x, y, L);
    QS_HIP(hipGetLastError());
}

}  // namespace qs
