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
// where u runs over the states the adjoint couplings reach from s.  Output
// rows are owned by one lane (or LPR lanes combined by a fixed shuffle
// tree): no atomics and a fixed summation order, so a result never depends
// on what else runs on the device.
//
// Two kernels:
//  * k_matvec_implicit -- no symmetry or spin inversion only.  The column is
//    the combinadic rank of the state (updated incrementally for two-site
//    hops), so there is no table at all.  One lane per row, a wave-uniform
//    loop over the flip groups (couplings read with scalar loads; for one
//    bond consecutive rows reach nearby columns, so gathers coalesce).
//  * k_matvec_stored -- translations / parity / inversions.  Each lane
//    collects the live flip groups of its share in a bit mask, then works on
//    them NB at a time: orbit minima from 32-bit funnel shifts, one probe of
//    the HBM hash table canonical state -> column, then the gathers of x and
//    1/sqrt(D).
// Both issue all loads of an NB-neighbour batch before consuming any, so
// memory latency is paid once per batch, and both renumber blocks so each
// XCD owns a contiguous row range (its own L2 stays warm).
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "qs_device.h"

namespace qs {

namespace {

constexpr int kMvBlock = 256;   // implicit-basis kernel
constexpr int kStBlock = 256;   // stored-basis kernel
constexpr int kMaxLdsGroups = 256;
constexpr int kMaxLdsTerms = 512;
constexpr int kNB = 8;          // neighbours per batch

struct MVArgs {
    int64_t N;
    int L, nup, zchar;
    uint64_t mask;
    GroupDev g;
    const void *reps;
    const double *inv_nrm;
    const uint64_t *hkeys;
    const uint32_t *hvals;
    uint64_t hmask;
    int hmax_probe;
    const uint64_t *binom;
    int ndiag;
    int noff;
    int ngroups;
    const double *user_diag;   // interleaved complex or null
    int user_diag_conj;
    double are, aim, bre, bim;
    int conj_x, conj_y;
    int beta_zero;
    double *dot_partial;   // optional: per-block FP64 partials of <x, y>
    // dynamic LDS layout (byte offsets)
    int off_binom, off_grp, off_term, off_red;
    // hop kernel
    const uint32_t *ptab;  // parent lookup table (stored mode)
    int colbits;
    uint32_t cfull;        // C(L, nup): spin inversion reverses the colex order
    int ndcls, ndrest;
    // SpinBasisGeneral: the characters of the parent table's group elements
    // (interleaved complex), else null
    const double *gchi;
};

// Character of the group element a parent-table entry names: (t, p) =
// divmod(elem, npoint) of the translation / point group, or the general
// group's element list.
__device__ __forceinline__ void elem_char(const MVArgs &A, int elem, double &cr, double &ci) {
    if (A.gchi) {
        cr = A.gchi[2 * elem];
        ci = A.gchi[2 * elem + 1];
        return;
    }
    const int bt = elem / A.g.npoint, bp = elem - bt * A.g.npoint;
    const double lr = A.g.lam_re[bt], li = A.g.lam_im[bt];
    const double pr = A.g.pch_re[bp], pi = A.g.pch_im[bp];
    cr = lr * pr - li * pi;
    ci = lr * pi + li * pr;
}

// Amplitude of one compiled term on state s: A * (-1)^popcount(s & smask),
// or zero when the conditions fail.
template <typename S, bool CC>
__device__ __forceinline__ void term_amp(const TermDev &t, S s, double &re, double &im) {
    if ((s & (S)t.cmask) != (S)t.cval) return;
    double sg = (popc(s & (S)t.smask) & 1) ? -1.0 : 1.0;
    re += sg * t.are;
    if (CC) im += sg * t.aim;
}

// Column of canonical state r in the hash table, or -1.  `e` is the first
// probe, already loaded by the caller.
template <typename S>
__device__ __forceinline__ int64_t hash_resolve(const MVArgs &A, S r, uint64_t h, uint64_t e) {
    for (int p = 0;; p++) {
        if (e == ~0ULL) return -1;
        if constexpr (sizeof(S) == 4) {
            if ((uint32_t)(e >> 32) == r) return (int64_t)(uint32_t)e;
        } else {
            if (e == (uint64_t)r) return (int64_t)A.hvals[h];
        }
        if (p + 1 >= A.hmax_probe) return -1;
        h = (h + 1) & A.hmask;
        e = A.hkeys[h];
    }
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

// Column of state u for the implicit bases; row is the rank of s and
// u = s ^ (one flip group).  -1 when u leaves the particle sector; cr
// receives the inversion character.
template <typename S, int MODE, bool LDSB>
__device__ __forceinline__ int64_t implicit_col(const MVArgs &A, S s, S u, int64_t row, int lo, int hi,
                                                const uint32_t *s_binom, double &cr) {
    cr = 1.0;
    if (A.nup < 0) {
        if (MODE == MODE_ZONLY && ((u >> (A.L - 1)) & 1)) {
            cr = (double)A.zchar;
            return (int64_t)(~u & (S)A.mask);
        }
        return (int64_t)u;
    }
    if (popc(u) != A.nup) return -1;
    if (MODE == MODE_ZONLY && ((u >> (A.L - 1)) & 1)) {
        cr = (double)A.zchar;
        const S r = ~u & (S)A.mask;
        return LDSB ? comb_rank<S, uint32_t>(r, s_binom) : comb_rank<S, uint64_t>(r, A.binom);
    }
    if (lo >= 0)
        return row + (LDSB ? rank_delta<S, uint32_t>(s, lo, hi, s_binom) : rank_delta<S, uint64_t>(s, lo, hi, A.binom));
    return LDSB ? comb_rank<S, uint32_t>(u, s_binom) : comb_rank<S, uint64_t>(u, A.binom);
}

template <int BLOCK>
__device__ __forceinline__ void write_dot_partial(const MVArgs &A, double *s_red, double dot_r, double dot_i) {
    s_red[threadIdx.x] = dot_r;
    s_red[BLOCK + threadIdx.x] = dot_i;
    __syncthreads();
    for (int st = BLOCK / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            s_red[threadIdx.x] += s_red[threadIdx.x + st];
            s_red[BLOCK + threadIdx.x] += s_red[BLOCK + threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        A.dot_partial[2 * blockIdx.x] = s_red[0];
        A.dot_partial[2 * blockIdx.x + 1] = s_red[BLOCK];
    }
}

// y[row] = beta*y + alpha*acc with optional conjugations; FP64 <x, y> partial
template <typename V>
__device__ __forceinline__ void finish_row(const MVArgs &A, V *y, int64_t row, double ar, double ai, double xr,
                                           double xi, double &dot_r, double &dot_i) {
    using T = VTraits<V>;
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

// One row of an implicit basis (no symmetry / spin inversion only).
template <typename S, typename V, bool CC, int MODE>
__device__ __forceinline__ void implicit_row(const MVArgs &A, const GroupOfTerms *__restrict__ grp,
                                             const TermDev *__restrict__ off, const TermDev *__restrict__ diag,
                                             const V *__restrict__ x, V *__restrict__ y, int64_t row,
                                             const uint32_t *s_binom, double &dot_r, double &dot_i) {
    using T = VTraits<V>;
    constexpr bool LDSB = sizeof(S) == 4;
    S s;
    if (A.nup >= 0) s = LDSB ? comb_unrank<S, uint32_t>(row, A.L, A.nup, s_binom)
                             : comb_unrank<S, uint64_t>(row, A.L, A.nup, A.binom);
    else s = (S)row;
    double xr, xi;
    T::load(x, row, xr, xi);
    if (A.conj_x) xi = -xi;
    double dr = 0, di = 0;
    for (int k = 0; k < A.ndiag; k++) term_amp<S, CC>(diag[k], s, dr, di);
    di = -di;   // conj(amp): the terms are the adjoint couplings
    if (A.user_diag) {
        dr += A.user_diag[2 * row];
        di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
    }
    double ar = dr * xr - di * xi;
    double ai = dr * xi + di * xr;
    for (int g0 = 0; g0 < A.ngroups; g0 += kNB) {
        int64_t col[kNB];
        double wr[kNB], wi[kNB];
#pragma unroll
        for (int j = 0; j < kNB; j++) {
            const int gi = g0 + j;
            col[j] = row;
            wr[j] = 0;
            wi[j] = 0;
            if (gi < A.ngroups) {
                const GroupOfTerms gr = grp[gi];
                double mr = 0, mi = 0;
                for (int k = gr.t0; k < gr.t1; k++) term_amp<S, CC>(off[k], s, mr, mi);
                if (mr != 0.0 || mi != 0.0) {
                    double cr;
                    const int64_t c = implicit_col<S, MODE, LDSB>(A, s, s ^ (S)gr.flip, row, gr.lo, gr.hi,
                                                                  s_binom, cr);
                    if (c >= 0) {
                        col[j] = c;
                        wr[j] = mr * cr;
                        wi[j] = -mi * cr;
                    }
                }
            }
        }
        double vr[kNB], vi[kNB];
#pragma unroll
        for (int j = 0; j < kNB; j++) T::load(x, col[j], vr[j], vi[j]);
#pragma unroll
        for (int j = 0; j < kNB; j++) {
            if (A.conj_x) vi[j] = -vi[j];
            ar += wr[j] * vr[j] - wi[j] * vi[j];
            ai += wr[j] * vi[j] + wi[j] * vr[j];
        }
    }
    finish_row<V>(A, y, row, ar, ai, xr, xi, dot_r, dot_i);
}

template <typename S, typename V, bool CC, int MODE>
__global__ __launch_bounds__(kMvBlock) void k_matvec_implicit(MVArgs A, const GroupOfTerms *__restrict__ grp,
                                                              const TermDev *__restrict__ off,
                                                              const TermDev *__restrict__ diag,
                                                              const V *__restrict__ x, V *__restrict__ y) {
    extern __shared__ __align__(16) char smem[];
    uint32_t *s_binom = (uint32_t *)(smem + A.off_binom);
    double *s_red = (double *)(smem + A.off_red);
    if (sizeof(S) == 4 && A.nup >= 0)
        for (int i = threadIdx.x; i < 33 * 65; i += kMvBlock) s_binom[i] = (uint32_t)A.binom[i];
    __syncthreads();
    const int vb = xcd_block(blockIdx.x, gridDim.x);
    double dot_r = 0, dot_i = 0;
    for (int64_t base = (int64_t)vb * kMvBlock; base < A.N; base += (int64_t)gridDim.x * kMvBlock) {
        const int64_t row = base + threadIdx.x;
        if (row < A.N) implicit_row<S, V, CC, MODE>(A, grp, off, diag, x, y, row, s_binom, dot_r, dot_i);
    }
    if (A.dot_partial) write_dot_partial<kMvBlock>(A, s_red, dot_r, dot_i);
}

// One tile of kStBlock / LPR rows starting at `base` (stored basis).
template <typename S, typename V, bool CC, bool TRACK, int LPR>
__device__ __forceinline__ void stored_tile(const MVArgs &A, const GroupOfTerms *G, const TermDev *TT,
                                            const TermDev *__restrict__ diag, const V *__restrict__ x,
                                            V *__restrict__ y, int64_t base, double &dot_r, double &dot_i) {
    using T = VTraits<V>;
    const S mask = (S)A.mask;
    const S *reps = (const S *)A.reps;
    const int sub = threadIdx.x % LPR;
    const int64_t row = base + threadIdx.x / LPR;
    const bool valid = row < A.N;
    double ar = 0, ai = 0, xr = 0, xi = 0;
    if (valid) {
        const S s = reps[row];
        const double own_scale = 1.0 / A.inv_nrm[row];
        if (sub == 0) {
            T::load(x, row, xr, xi);
            if (A.conj_x) xi = -xi;
            double dr = 0, di = 0;
            for (int k = 0; k < A.ndiag; k++) term_amp<S, CC>(diag[k], s, dr, di);
            di = -di;
            if (A.user_diag) {
                dr += A.user_diag[2 * row];
                di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
            }
            ar = dr * xr - di * xi;
            ai = dr * xi + di * xr;
        }
        for (int g0 = sub; g0 < A.ngroups; g0 += 64 * LPR) {
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
                // a batch of up to kNB live neighbours
                S u[kNB];
                double mr[kNB], mi[kNB];
                bool act[kNB];
#pragma unroll
                for (int j = 0; j < kNB; j++) {
                    act[j] = live != 0;
                    mr[j] = 0;
                    mi[j] = 0;
                    u[j] = s;
                    if (act[j]) {
                        const int jj = __builtin_ctzll(live);
                        live &= live - 1;
                        const GroupOfTerms gr = G[g0 + jj * LPR];
                        for (int k = gr.t0; k < gr.t1; k++) term_amp<S, CC>(TT[k], s, mr[j], mi[j]);
                        mi[j] = -mi[j];
                        u[j] = s ^ (S)gr.flip;
                    }
                }
                S r[kNB];
                int bt[kNB], bp[kNB];
                if constexpr (sizeof(S) == 4) {
                    canonical32_batch<kNB, TRACK>(A.g, u, r, bt, bp);
                } else {
#pragma unroll
                    for (int j = 0; j < kNB; j++) r[j] = canonical<S, TRACK>(A.g, u[j], mask, bt[j], bp[j]);
                }
                // first probe of every neighbour in flight together
                uint64_t h[kNB], e[kNB];
#pragma unroll
                for (int j = 0; j < kNB; j++) {
                    h[j] = state_hash((uint64_t)r[j]) & A.hmask;
                    e[j] = A.hkeys[h[j]];
                }
                int64_t col[kNB];
#pragma unroll
                for (int j = 0; j < kNB; j++) {
                    col[j] = act[j] ? hash_resolve<S>(A, r[j], h[j], e[j]) : -1;
                    if (col[j] < 0) act[j] = false;
                }
                double f[kNB], vr[kNB], vi[kNB];
#pragma unroll
                for (int j = 0; j < kNB; j++) {
                    const int64_t c = act[j] ? col[j] : row;
                    f[j] = A.inv_nrm[c];
                    T::load(x, c, vr[j], vi[j]);
                }
#pragma unroll
                for (int j = 0; j < kNB; j++) {
                    if (!act[j]) continue;
                    double cr = 1.0, ci = 0.0;
                    if (TRACK) {
                        const double lr = A.g.lam_re[bt[j]], li = A.g.lam_im[bt[j]];
                        const double pr = A.g.pch_re[bp[j]], pi = A.g.pch_im[bp[j]];
                        cr = lr * pr - li * pi;
                        ci = lr * pi + li * pr;
                    }
                    const double fac = f[j] * own_scale;
                    cr *= fac;
                    ci *= fac;
                    const double wr = mr[j] * cr - mi[j] * ci;
                    const double wi = mr[j] * ci + mi[j] * cr;
                    double v_i = vi[j];
                    if (A.conj_x) v_i = -v_i;
                    ar += wr * vr[j] - wi * v_i;
                    ai += wr * v_i + wi * vr[j];
                }
            }
        }
    }
    // fixed-order combination of the LPR partial sums
#pragma unroll
    for (int o = LPR / 2; o >= 1; o >>= 1) {
        ar += __shfl_xor(ar, o);
        ai += __shfl_xor(ai, o);
    }
    if (valid && sub == 0) finish_row<V>(A, y, row, ar, ai, xr, xi, dot_r, dot_i);
}

template <typename S, typename V, bool CC, bool TRACK, int LPR, bool TLDS>
__global__ __launch_bounds__(kStBlock) void k_matvec_stored(MVArgs A, const GroupOfTerms *__restrict__ grp,
                                                            const TermDev *__restrict__ off,
                                                            const TermDev *__restrict__ diag,
                                                            const V *__restrict__ x, V *__restrict__ y) {
    extern __shared__ __align__(16) char smem[];
    GroupOfTerms *s_grp = (GroupOfTerms *)(smem + A.off_grp);
    TermDev *s_term = (TermDev *)(smem + A.off_term);
    double *s_red = (double *)(smem + A.off_red);
    if (TLDS) {
        for (int i = threadIdx.x; i < A.ngroups; i += kStBlock) s_grp[i] = grp[i];
        for (int i = threadIdx.x; i < A.noff; i += kStBlock) s_term[i] = off[i];
    }
    __syncthreads();
    const GroupOfTerms *G = TLDS ? s_grp : grp;
    const TermDev *TT = TLDS ? s_term : off;
    constexpr int kRows = kStBlock / LPR;
    const int vb = xcd_block(blockIdx.x, gridDim.x);
    double dot_r = 0, dot_i = 0;
    for (int64_t base = (int64_t)vb * kRows; base < A.N; base += (int64_t)gridDim.x * kRows)
        stored_tile<S, V, CC, TRACK, LPR>(A, G, TT, diag, x, y, base, dot_r, dot_i);
    if (A.dot_partial) write_dot_partial<kStBlock>(A, s_red, dot_r, dot_i);
}

// ----------------------------------------------------------- wide bases --
// Bases of up to 256 sites (states as four words).  LPR = 8 lanes per row,
// lane `sub` taking the flip groups g = sub (mod 8), as in the stored
// kernel: per live group, the flipped state's orbit minimum under the group
// (the first element reaching it, with its character), its column by binary
// search over the ascending orbit minima, and the weight m chi sqrt(D_r / D_c).
struct WideArgs {
    const uint64_t *reps;
    WGroup g;
    const WGroupDev *grp;
    const WTermDev *off, *diag;
};

template <bool CC>
__device__ __forceinline__ void wterm_amp(const WTermDev &t, const W4 &s, double &re, double &im) {
    int par = 0;
    for (int k = 0; k < 4; k++) {
        if ((s.w[k] & t.cmask[k]) != t.cval[k]) return;
        par ^= __builtin_popcountll(s.w[k] & t.smask[k]) & 1;
    }
    const double sg = par ? -1.0 : 1.0;
    re += sg * t.are;
    if (CC) im += sg * t.aim;
}

constexpr int kWideLPR = 8;

template <typename V, bool CC, bool TRACK>
__global__ __launch_bounds__(kStBlock) void k_matvec_wide(MVArgs A, WideArgs W, const V *__restrict__ x,
                                                          V *__restrict__ y) {
    using T = VTraits<V>;
    __shared__ double s_red[2 * kStBlock];
    constexpr int kRows = kStBlock / kWideLPR;
    const int sub = threadIdx.x % kWideLPR;
    double dot_r = 0, dot_i = 0;
    for (int64_t base = (int64_t)blockIdx.x * kRows; base < A.N; base += (int64_t)gridDim.x * kRows) {
        const int64_t row = base + threadIdx.x / kWideLPR;
        const bool valid = row < A.N;
        double ar = 0, ai = 0, xr = 0, xi = 0;
        if (valid) {
            const W4 s = w4_at(W.reps, row);
            const double own_scale = 1.0 / A.inv_nrm[row];
            if (sub == 0) {
                T::load(x, row, xr, xi);
                if (A.conj_x) xi = -xi;
                double dr = 0, di = 0;
                for (int k = 0; k < A.ndiag; k++) wterm_amp<CC>(W.diag[k], s, dr, di);
                di = -di;
                if (A.user_diag) {
                    dr += A.user_diag[2 * row];
                    di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
                }
                ar = dr * xr - di * xi;
                ai = dr * xi + di * xr;
            }
            for (int gi = sub; gi < A.ngroups; gi += kWideLPR) {
                const WGroupDev gr = W.grp[gi];
                double mr = 0, mi = 0;
                for (int k = gr.t0; k < gr.t1; k++) wterm_amp<CC>(W.off[k], s, mr, mi);
                if (mr == 0.0 && mi == 0.0) continue;
                mi = -mi;
                W4 u;
                for (int k = 0; k < 4; k++) u.w[k] = s.w[k] ^ gr.flip[k];
                W4 best = u;
                int be = 0;
                for (int e = 1; e < W.g.n; e++) {
                    const W4 w = wg_apply(W.g, e, u);
                    if (w4_less(w, best)) {
                        best = w;
                        be = e;
                    }
                }
                const int64_t col = w4_find(W.reps, A.N, best);
                if (col < 0) continue;
                double cr = 1.0, ci = 0.0;
                if (TRACK) {
                    cr = W.g.chi[2 * be];
                    ci = W.g.chi[2 * be + 1];
                }
                const double fac = A.inv_nrm[col] * own_scale;
                cr *= fac;
                ci *= fac;
                const double wr = mr * cr - mi * ci;
                const double wi = mr * ci + mi * cr;
                double vr, vi;
                T::load(x, col, vr, vi);
                if (A.conj_x) vi = -vi;
                ar += wr * vr - wi * vi;
                ai += wr * vi + wi * vr;
            }
        }
#pragma unroll
        for (int o = kWideLPR / 2; o >= 1; o >>= 1) {
            ar += __shfl_xor(ar, o);
            ai += __shfl_xor(ai, o);
        }
        if (valid && sub == 0) finish_row<V>(A, y, row, ar, ai, xr, xi, dot_r, dot_i);
    }
    if (A.dot_partial) write_dot_partial<kStBlock>(A, s_red, dot_r, dot_i);
}

// ------------------------------------------------------------ hop kernel --
// Every flip group is a two-bit hop whose amplitude is a function of its two
// bits (HopDev), L <= 32 and all indices fit 31 bits.  One lane per row and a
// wave-uniform loop over the groups (group data in scalar registers).  The
// neighbour's column comes from 32-bit colex arithmetic alone:
//  * CM_FULL  -- the column is the rank itself (or the state without nup);
//  * CM_ZONLY -- canonical = top bit clear; an inverted neighbour's column
//                is C(L, nup) - 1 - rank, since inversion reverses the
//                colex order of the half-filled words;
//  * CM_PTAB  -- the rank indexes the parent lookup table, which holds the
//                column and the group element reaching the orbit minimum.
enum { CM_FULL = 0, CM_ZONLY = 1, CM_PTAB = 2 };

// diagonal of state s from the sign classes and the remaining terms
template <bool CC>
__device__ __forceinline__ void diag_value(const MVArgs &A, const DiagClass *__restrict__ dcls,
                                           const TermDev *__restrict__ drest, uint32_t s, double &dr, double &di) {
    const uint32_t mask = (uint32_t)A.mask;
    for (int k = 0; k < A.ndcls; k++) {
        const DiagClass c = dcls[k];
        const uint32_t xs = c.d ? (s ^ (((s >> c.d) | (s << (A.L - c.d))) & mask)) : s;
        const double f = (double)(c.n - 2 * popc(xs & (uint32_t)c.mask));
        dr += f * c.re;
        if (CC) di += f * c.im;
    }
    for (int k = 0; k < A.ndrest; k++) term_amp<uint32_t, CC>(drest[k], s, dr, di);
}

// Own state of a row: the state, its parent rank (CM_PTAB) or row index.
template <int CM, bool PC>
__device__ __forceinline__ void hop_state(const MVArgs &A, uint32_t row, const uint32_t *s_binom, uint32_t &s,
                                          uint32_t &rs) {
    const uint32_t *reps = (const uint32_t *)A.reps;
    if (CM == CM_PTAB) {
        s = reps[row];
        rs = PC ? comb_rank32(s, s_binom) : s;
    } else {
        s = reps ? reps[row] : (PC ? comb_unrank<uint32_t, uint32_t>(row, A.L, A.nup, s_binom) : row);
        rs = row;
    }
}

// Neighbour of state s (parent rank / column rs) under hop group h: whether
// the group acts on s, the target (column for the implicit modes, parent
// rank for the parent table) and the amplitude with the inversion character.
// The same enumeration as hop_groups, which keeps its copy inline: sharing
// this function there costs the matvec registers and occupancy.
template <bool CC, int CM, bool PC>
__device__ __forceinline__ void hop_target(const MVArgs &A, const HopDev &h, uint32_t s, uint32_t rs,
                                           const uint32_t *s_binom, bool &live, uint32_t &c, double &m_r,
                                           double &m_i) {
    const uint32_t mask = (uint32_t)A.mask;
    const uint32_t blo = (s >> h.lo) & 1, bhi = (s >> h.hi) & 1;
    live = (h.nzmask >> (blo | (bhi << 1))) & 1;
    m_i = 0;
    if (PC) {
        if (h.sym) {
            m_r = h.re[1];
            if (CC) m_i = h.im[1];
        } else {
            m_r = blo ? h.re[1] : h.re[2];
            if (CC) m_i = blo ? h.im[1] : h.im[2];
        }
    } else {
        m_r = bhi ? (blo ? h.re[3] : h.re[2]) : (blo ? h.re[1] : h.re[0]);
        if (CC) m_i = bhi ? (blo ? h.im[3] : h.im[2]) : (blo ? h.im[1] : h.im[0]);
    }
    if (PC) {
        // colex rank change of the hop (bits lo and hi differ)
        const int cl = popc(s & h.lowmask);
        uint32_t bw = s & h.between;
        uint32_t d = s_binom[h.hi65 + cl + popc(bw) + 1] - s_binom[h.lo65 + cl + 1];
        if (h.between) {
            int l = 1;
            while (bw) {
                const int q = ctz(bw);
                d += s_binom[q * 65 + cl + l] - s_binom[q * 65 + cl + l + 1];
                l++;
                bw &= bw - 1;
            }
        }
        c = rs + (blo ? d : 0u - d);
    } else {
        c = s ^ ((1u << h.lo) | (1u << h.hi));
    }
    if (CM == CM_ZONLY && h.top) {
        // s has bit L-1 clear, so the hop sets it: the column is
        // the inverted state's, with the inversion character
        c = PC ? A.cfull - 1 - c : (~c & mask);
        if (A.zchar < 0) {
            m_r = -m_r;
            m_i = -m_i;
        }
    }
}

// Element i of base through a 32-bit byte offset (the caller guarantees
// i * sizeof(T) < 2^32): the load takes the scalar-base + vector-offset form
// and needs no 64-bit address arithmetic.
template <typename T>
__device__ __forceinline__ const T *at32(const T *base, uint32_t i) {
    return (const T *)((const char *)base + (uint32_t)(i * (uint32_t)sizeof(T)));
}

// The element at byte offset `b` from base (b < 2^32): scalar base plus a
// 32-bit vector offset, as at32 but with the offset already in bytes.
template <typename T>
__device__ __forceinline__ const T *atb(const T *base, uint32_t b) {
    return (const T *)((const char *)base + b);
}

// FP64 form of a vector type (the pre-scaled x of the parent-table kernel)
template <typename V> struct Wide { using type = double; };
template <> struct Wide<cplx_f> { using type = cplx_d; };
template <> struct Wide<cplx_d> { using type = cplx_d; };

// xs = x * (1/sqrt(D)) column by column, in FP64: the parent-table hop
// kernel then gathers one value per neighbour instead of x and 1/sqrt(D).
template <typename V>
__global__ __launch_bounds__(256) void k_prescale(const V *__restrict__ x, const double *__restrict__ inv_nrm,
                                                  typename Wide<V>::type *__restrict__ xs, int64_t n) {
    using W = typename Wide<V>::type;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double re, im;
        VTraits<V>::load(x, i, re, im);
        const double f = inv_nrm[i];
        VTraits<W>::store(xs, i, re * f, im * f);
    }
}

// Adds the flip groups g = g_first, g_first + WS, ... of row `row` (own
// state s, rank rs) to (ar, ai), kNB neighbours per batch.  PS: neighbours
// are gathered from the pre-scaled xs (parent-table mode only), the parent
// table and xs through 32-bit byte offsets (the host checks they fit).
template <typename V, bool CC, int CM, bool TRACK, bool PC, int WS, bool PS>
__device__ __forceinline__ void hop_groups(const MVArgs &A, const HopDev *__restrict__ hops,
                                           const V *__restrict__ x, const typename Wide<V>::type *__restrict__ xs,
                                           uint32_t row, uint32_t s, uint32_t rs, int g_first,
                                           const uint32_t *s_binom, double &ar, double &ai) {
    using T = VTraits<V>;
    constexpr bool VC = T::cplx;
    constexpr int NB = kNB;
    const uint32_t mask = (uint32_t)A.mask;
    double own_scale = 1.0;
    if (CM == CM_PTAB) own_scale = 1.0 / A.inv_nrm[row];
    const uint32_t colmask = A.colbits >= 32 ? 0xFFFFFFFFu : ((1u << A.colbits) - 1);
    for (int g0 = g_first; g0 < A.ngroups; g0 += NB * WS) {
        uint32_t col[NB];
        double mr[NB], mi[NB];
        int elem[NB];
        uint32_t livebits = 0;   // neighbour j acts on this row
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const int gi = g0 + j * WS;
            col[j] = CM == CM_PTAB ? rs : row;
            mr[j] = 0;
            mi[j] = 0;
            elem[j] = 0;
            if (gi < A.ngroups) {
                // group constants are wave-uniform: scalar loads and branches
                const HopDev h = hops[gi];
                const uint32_t blo = (s >> h.lo) & 1, bhi = (s >> h.hi) & 1;
                const bool live = (h.nzmask >> (blo | (bhi << 1))) & 1;
                double m_r, m_i = 0;
                if (PC) {
                    if (h.sym) {
                        m_r = h.re[1];
                        if (CC) m_i = h.im[1];
                    } else {
                        m_r = blo ? h.re[1] : h.re[2];
                        if (CC) m_i = blo ? h.im[1] : h.im[2];
                    }
                } else {
                    m_r = bhi ? (blo ? h.re[3] : h.re[2]) : (blo ? h.re[1] : h.re[0]);
                    if (CC) m_i = bhi ? (blo ? h.im[3] : h.im[2]) : (blo ? h.im[1] : h.im[0]);
                }
                uint32_t c;
                if (PC) {
                    // colex rank change of the hop (bits lo and hi differ)
                    const int cl = popc(s & h.lowmask);
                    uint32_t bw = s & h.between;
                    uint32_t d = s_binom[h.hi65 + cl + popc(bw) + 1] - s_binom[h.lo65 + cl + 1];
                    if (h.between) {
                        int l = 1;
                        while (bw) {
                            const int q = ctz(bw);
                            d += s_binom[q * 65 + cl + l] - s_binom[q * 65 + cl + l + 1];
                            l++;
                            bw &= bw - 1;
                        }
                    }
                    c = rs + (blo ? d : 0u - d);
                } else {
                    c = s ^ ((1u << h.lo) | (1u << h.hi));
                }
                if (CM == CM_ZONLY && h.top) {
                    // s has bit L-1 clear, so the hop sets it: the column is
                    // the inverted state's, with the inversion character
                    c = PC ? A.cfull - 1 - c : (~c & mask);
                    if (A.zchar < 0) {
                        m_r = -m_r;
                        m_i = -m_i;
                    }
                }
                // branch-free: dead neighbours keep col = own row, weight 0
                col[j] = live ? c : col[j];
                mr[j] = live ? m_r : 0.0;
                mi[j] = live ? -m_i : 0.0;
                livebits |= (uint32_t)live << j;
            }
        }
        double f[NB];
        if (CM == CM_PTAB) {
            // dead neighbours issue no parent-table request (their lanes are
            // masked off the gather) and fall back to the own row below
            uint32_t e[NB];
#pragma unroll
            for (int j = 0; j < NB; j++)
                e[j] = ((livebits >> j) & 1) ? (PS ? *at32(A.ptab, col[j]) : A.ptab[col[j]]) : 0xFFFFFFFFu;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                if (e[j] == 0xFFFFFFFFu) {
                    mr[j] = 0;
                    mi[j] = 0;
                    col[j] = row;
                } else {
                    col[j] = e[j] & colmask;
                    // a batched launch may run a basis without characters
                    // (no element bits, colbits = 32) in the tracking kernel
                    elem[j] = (TRACK && A.colbits < 32) ? (int)(e[j] >> A.colbits) : 0;
                }
            }
#pragma unroll
            for (int j = 0; j < NB; j++) f[j] = PS ? 1.0 : A.inv_nrm[col[j]];
        }
        double vr[NB], vi[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (PS) VTraits<typename Wide<V>::type>::load(at32(xs, col[j]), 0, vr[j], vi[j]);
            else T::load(x, col[j], vr[j], vi[j]);
        }
        // real couplings without characters: the weight is mr * fac, the
        // same value the complex form gives
#pragma unroll
        for (int j = 0; j < NB; j++) {
            double wr = mr[j], wi = CC ? mi[j] : 0.0;
            if (CM == CM_PTAB) {
                const double fac = f[j] * own_scale;
                if (TRACK) {
                    double cr, ci;
                    elem_char(A, elem[j], cr, ci);
                    cr *= fac;
                    ci *= fac;
                    wr = CC ? mr[j] * cr - mi[j] * ci : mr[j] * cr;
                    wi = CC ? mr[j] * ci + mi[j] * cr : mr[j] * ci;
                } else {
                    wr = mr[j] * fac;
                    if (CC) wi = mi[j] * fac;
                }
            }
            if (VC) {
                // one form for every weight: a job gives the same bits in the
                // character-tracking kernel a batch may run it in
                double v_i = vi[j];
                if (A.conj_x) v_i = -v_i;
                ar += wr * vr[j] - wi * v_i;
                ai += wr * v_i + wi * vr[j];
            } else {
                ar += wr * vr[j];
            }
        }
    }
}

// Block b of nb (kMvBlock threads = kMvBlock / 64 waves).  WS waves share
// each 64-row tile: wave `sub` takes the flip groups g = sub (mod WS) (so
// group data stays wave-uniform) and the partial sums meet in LDS in a fixed
// order.  (b, nb) are explicit so a batched launch gives every job the
// partition of its own single launch.
template <typename V, bool CC, int CM, bool TRACK, bool PC, int WS, bool PS = false>
__device__ __forceinline__ void hop_body(const MVArgs &A, const HopDev *__restrict__ hops,
                                         const DiagClass *__restrict__ dcls, const TermDev *__restrict__ drest,
                                         const V *__restrict__ x, const typename Wide<V>::type *__restrict__ xs,
                                         V *__restrict__ y, int b, int nb, char *smem) {
    using T = VTraits<V>;
    constexpr bool VC = T::cplx;
    constexpr int kRows = kMvBlock / WS;
    uint32_t *s_binom = (uint32_t *)(smem + A.off_binom);
    double *s_red = (double *)(smem + A.off_red);
    __shared__ double s_part[WS > 1 ? 2 * kMvBlock : 1];
    if (PC)
        for (int i = threadIdx.x; i < 33 * 65; i += kMvBlock) s_binom[i] = (uint32_t)A.binom[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // wave-uniform as far as the compiler can tell, so the flip-group
    // constants of hop_groups stay scalar loads
    const int sub = __builtin_amdgcn_readfirstlane(wave % WS), rw = wave / WS;
    const int vb = xcd_block(b, nb);
    const uint32_t n = (uint32_t)A.N;
    double dot_r = 0, dot_i = 0;
    for (uint32_t base = (uint32_t)vb * kRows; base < n; base += (uint32_t)nb * kRows) {
        const uint32_t row = base + rw * 64 + lane;
        double ar = 0, ai = 0, xr = 0, xi = 0;
        if (row < n) {
            uint32_t s, rs;
            hop_state<CM, PC>(A, row, s_binom, s, rs);
            if (sub == 0) {
                T::load(x, row, xr, xi);
                if (A.conj_x) xi = -xi;
                double dr = 0, di = 0;
                diag_value<CC>(A, dcls, drest, s, dr, di);
                di = -di;   // conj(amp): the terms are the adjoint couplings
                if (A.user_diag) {
                    dr += A.user_diag[2 * row];
                    di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
                }
                if (VC) {
                    ar = dr * xr - di * xi;
                    ai = dr * xi + di * xr;
                } else {
                    ar = dr * xr;
                }
            }
            hop_groups<V, CC, CM, TRACK, PC, WS, PS>(A, hops, x, xs, row, s, rs, sub, s_binom, ar, ai);
        }
        if (WS > 1) {
            s_part[threadIdx.x] = ar;
            s_part[kMvBlock + threadIdx.x] = ai;
            __syncthreads();
            if (sub == 0) {
#pragma unroll
                for (int k = 1; k < WS; k++) {
                    ar += s_part[threadIdx.x + 64 * k];
                    ai += s_part[kMvBlock + threadIdx.x + 64 * k];
                }
            }
            __syncthreads();
        }
        if (sub == 0 && row < n) finish_row<V>(A, y, row, ar, ai, xr, xi, dot_r, dot_i);
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
            A.dot_partial[2 * b] = s_red[0];
            A.dot_partial[2 * b + 1] = s_red[kMvBlock];
        }
    }
}

template <typename V, bool CC, int CM, bool TRACK, bool PC, int WS, bool PS>
__global__ __launch_bounds__(kMvBlock) void k_matvec_hop(MVArgs A, const HopDev *__restrict__ hops,
                                                         const DiagClass *__restrict__ dcls,
                                                         const TermDev *__restrict__ drest, const V *__restrict__ x,
                                                         const typename Wide<V>::type *__restrict__ xs,
                                                         V *__restrict__ y) {
    extern __shared__ __align__(16) char smem[];
    hop_body<V, CC, CM, TRACK, PC, WS, PS>(A, hops, dcls, drest, x, xs, y, blockIdx.x, gridDim.x, smem);
}

// Batched hop matvec: job j owns blocks [b0_j, b0_j + nb_j) of the grid and
// runs them exactly as its own launch of nb_j blocks would.
struct MVJob {
    const MVArgs *A;   // device copy
    const HopDev *hops;
    const DiagClass *dcls;
    const TermDev *drest;
    const void *x;
    void *y;
    int b0, nb;
};
constexpr int kMaxMVJobs = 32;
struct MVJobs {
    int njobs;
    MVJob job[kMaxMVJobs];
};

template <typename V, bool CC, int CM, bool TRACK, bool PC>
__global__ __launch_bounds__(kMvBlock) void kb_matvec_hop(MVJobs J) {
    extern __shared__ __align__(16) char smem[];
    int lo = 0, hi = J.njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)blockIdx.x >= J.job[mid].b0) lo = mid;
        else hi = mid - 1;
    }
    const MVJob &q = J.job[lo];
    hop_body<V, CC, CM, TRACK, PC, 1>(*q.A, q.hops, q.dcls, q.drest, (const V *)q.x, nullptr, (V *)q.y,
                                      (int)blockIdx.x - q.b0, q.nb, smem);
}

// ------------------------------------------------ assembled sector matrix --
// The sector matrix M = P'HP of a hop-eligible operator, assembled once in
// sliced ELL form (SELL-64): rows in slices of 64; slice q is padded to the
// length w_q of its longest row and stored column-major (entry k of row
// 64q + i at sptr[q] + 64k + i), so a wavefront's loads coalesce.  A row's
// entries are its diagonal, then every live hop group's neighbour in group
// order, with the weights the pull kernel applies; padding points at the row
// itself with weight 0.  Duplicate columns stay separate entries (the
// product sums them); qs_op_csr merges them for export.
constexpr int kSellBlock = 256;

// one row's entries: emit(col, w_re, w_im)
template <bool CC, int CM, bool TRACK, bool PC, typename F>
__device__ __forceinline__ void row_entries(const MVArgs &A, const HopDev *__restrict__ hops,
                                            const DiagClass *__restrict__ dcls, const TermDev *__restrict__ drest,
                                            uint32_t row, const uint32_t *s_binom, F &&emit) {
    uint32_t s, rs;
    hop_state<CM, PC>(A, row, s_binom, s, rs);
    double dr = 0, di = 0;
    diag_value<CC>(A, dcls, drest, s, dr, di);
    di = -di;   // conj(amp): the terms are the adjoint couplings
    if (A.user_diag) {
        dr += A.user_diag[2 * row];
        di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
    }
    emit(row, dr, di);
    double own_scale = 1.0;
    if (CM == CM_PTAB) own_scale = 1.0 / A.inv_nrm[row];
    const uint32_t colmask = A.colbits >= 32 ? 0xFFFFFFFFu : ((1u << A.colbits) - 1);
    for (int g = 0; g < A.ngroups; g++) {
        const HopDev h = hops[g];
        bool live;
        uint32_t c;
        double m_r, m_i;
        hop_target<CC, CM, PC>(A, h, s, rs, s_binom, live, c, m_r, m_i);
        if (!live) continue;
        double wr = m_r, wi = -m_i;
        uint32_t col = c;
        if (CM == CM_PTAB) {
            const uint32_t e = A.ptab[c];
            if (e == 0xFFFFFFFFu) continue;
            col = e & colmask;
            double cr = 1.0, ci = 0.0;
            if (TRACK && A.colbits < 32) elem_char(A, (int)(e >> A.colbits), cr, ci);
            const double fac = A.inv_nrm[col] * own_scale;
            cr *= fac;
            ci *= fac;
            wr = m_r * cr + m_i * ci;
            wi = m_r * ci - m_i * cr;
        }
        emit(col, wr, wi);
    }
}

template <bool CC, int CM, bool TRACK, bool PC>
__global__ __launch_bounds__(kSellBlock) void k_sell_count(MVArgs A, const HopDev *__restrict__ hops,
                                                           const DiagClass *__restrict__ dcls,
                                                           const TermDev *__restrict__ drest, uint32_t *cnt) {
    extern __shared__ __align__(16) char smem[];
    uint32_t *s_binom = (uint32_t *)(smem + A.off_binom);
    if (PC)
        for (int i = threadIdx.x; i < 33 * 65; i += kSellBlock) s_binom[i] = (uint32_t)A.binom[i];
    __syncthreads();
    const uint32_t row = blockIdx.x * kSellBlock + threadIdx.x;
    if (row >= (uint32_t)A.N) return;
    uint32_t n = 0;
    row_entries<CC, CM, TRACK, PC>(A, hops, dcls, drest, row, s_binom, [&](uint32_t, double, double) { n++; });
    cnt[row] = n;
}

template <bool CC, int CM, bool TRACK, bool PC>
__global__ __launch_bounds__(kSellBlock) void k_sell_fill(MVArgs A, const HopDev *__restrict__ hops,
                                                          const DiagClass *__restrict__ dcls,
                                                          const TermDev *__restrict__ drest,
                                                          const uint64_t *__restrict__ sptr,
                                                          const uint32_t *__restrict__ width, uint32_t *cols,
                                                          double *wre, double *wim) {
    extern __shared__ __align__(16) char smem[];
    uint32_t *s_binom = (uint32_t *)(smem + A.off_binom);
    if (PC)
        for (int i = threadIdx.x; i < 33 * 65; i += kSellBlock) s_binom[i] = (uint32_t)A.binom[i];
    __syncthreads();
    const uint32_t row = blockIdx.x * kSellBlock + threadIdx.x;
    if (row >= (uint32_t)A.N) return;
    const uint32_t q = row >> 6, i = row & 63;
    const uint64_t base = sptr[q] + i;
    uint32_t k = 0;
    row_entries<CC, CM, TRACK, PC>(A, hops, dcls, drest, row, s_binom, [&](uint32_t col, double wr, double wi) {
        cols[base + 64 * (uint64_t)k] = col;
        wre[base + 64 * (uint64_t)k] = wr;
        if (wim) wim[base + 64 * (uint64_t)k] = wi;
        k++;
    });
    for (; k < width[q]; k++) {
        cols[base + 64 * (uint64_t)k] = row;
        wre[base + 64 * (uint64_t)k] = 0.0;
        if (wim) wim[base + 64 * (uint64_t)k] = 0.0;
    }
}

// y = beta*y + alpha*M x from the assembled rows (WC: complex weights)
template <typename V, bool WC>
__global__ __launch_bounds__(kSellBlock) void k_sell_spmv(MVArgs A, const uint64_t *__restrict__ sptr,
                                                          const uint32_t *__restrict__ width,
                                                          const uint32_t *__restrict__ cols, SellW W,
                                                          const V *__restrict__ x, V *__restrict__ y) {
    const int64_t row = (int64_t)blockIdx.x * kSellBlock + threadIdx.x;
    if (row >= A.N) return;
    const uint64_t q = (uint64_t)row >> 6;
    double ar = 0, ai = 0;
    if (!WC) W.im = nullptr;
    sell_row<V>(cols, W, sptr[q] + (row & 63), width[q], x, A.conj_x, ar, ai);
    double dot_r = 0, dot_i = 0;
    finish_row<V>(A, y, row, ar, ai, 0.0, 0.0, dot_r, dot_i);
}

// Batched products from assembled operators: job j owns blocks [b0_j, b0_j +
// nb_j) and runs exactly its single launch's rows (one lane per row), so its
// result is bitwise the same alone and in any batch.
struct SellJob {
    const MVArgs *A;   // device copy
    const uint64_t *sptr;
    const uint32_t *width, *cols;
    SellW wt;
    const void *x;
    void *y;
    int b0, nb;
    const void *z;     // optional: y += a z, a from zc (CoefMode zmode)
    const double *zc;
    int zmode;
    double *dot;       // optional: per-block partials of <x, y>
};
constexpr int kSellJobsInline = 24;   // jobs passed as kernel arguments (within 4 KB)
struct SellJobs {
    int njobs;
    SellJob job[kSellJobsInline];
};

// block b of job q: its kSellBlock rows
template <typename V>
__device__ __forceinline__ void sell_job_block(const MVArgs &A, const SellJob &q, int b) {
    using T = VTraits<V>;
    const int64_t row = (int64_t)b * kSellBlock + threadIdx.x;
    if (!q.dot) {
        if (row >= A.N) return;
        double ar = 0, ai = 0;
        sell_row<V>(q.cols, q.wt, q.sptr[(uint64_t)row >> 6] + (row & 63), q.width[(uint64_t)row >> 6],
                    (const V *)q.x, A.conj_x, ar, ai);
        double dot_r = 0, dot_i = 0;
        finish_row<V>(A, (V *)q.y, row, ar, ai, 0.0, 0.0, dot_r, dot_i);
        return;
    }
    // the Lanczos form: y = M x + a z (the axpby's operation order), then
    // this block's partial of <x, y> (the Gram product's per-row terms)
    __shared__ double red[2 * (kSellBlock / 64)];
    double dr = 0, di = 0;
    if (row < A.N) {
        double ar = 0, ai = 0;
        sell_row<V>(q.cols, q.wt, q.sptr[(uint64_t)row >> 6] + (row & 63), q.width[(uint64_t)row >> 6],
                    (const V *)q.x, A.conj_x, ar, ai);
        double dot_r = 0, dot_i = 0;
        finish_row<V>(A, (V *)q.y, row, ar, ai, 0.0, 0.0, dot_r, dot_i);
        double yr, yi, xr, xi;
        T::load((const V *)q.y, row, yr, yi);
        if (q.z) {
            double cr = 0, ci = 0, zr, zi;
            dev_coef(q.zc, q.zmode, cr, ci);
            T::load((const V *)q.z, row, zr, zi);
            const double orr = cr * zr - ci * zi + (1.0 * yr - 0.0 * yi);
            const double oi = cr * zi + ci * zr + (1.0 * yi + 0.0 * yr);
            T::store((V *)q.y, row, orr, oi);
            T::load((const V *)q.y, row, yr, yi);
        }
        T::load((const V *)q.x, row, xr, xi);
        dr = xr * yr + xi * yi;
        di = xr * yi - xi * yr;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        dr += __shfl_xor(dr, o, 64);
        di += __shfl_xor(di, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w] = dr;
        red[kSellBlock / 64 + w] = di;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        q.dot[2 * b] = (red[0] + red[1]) + (red[2] + red[3]);
        q.dot[2 * b + 1] = (red[4] + red[5]) + (red[6] + red[7]);
    }
}

template <typename V>
__global__ __launch_bounds__(kSellBlock) void kb_sell_spmv(SellJobs J) {
    int lo = 0, hi = J.njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)blockIdx.x >= J.job[mid].b0) lo = mid;
        else hi = mid - 1;
    }
    sell_job_block<V>(*J.job[lo].A, J.job[lo], (int)blockIdx.x - J.job[lo].b0);
}

// One operator's Lanczos product on its stored rows (the job's A unused:
// the arguments come by value)
template <typename V>
__global__ __launch_bounds__(kSellBlock) void k_sell_lanczos(MVArgs A, SellJob q) {
    sell_job_block<V>(A, q, (int)blockIdx.x);
}

// The same over a device job table: job j owns blocks [j nbmax, (j+1) nbmax)
// and uses the first tab[j].nb of them.
template <typename V>
__global__ __launch_bounds__(kSellBlock) void kb_sell_tab(const SellJob *__restrict__ tab, int nbmax) {
    const int j = (int)blockIdx.x / nbmax, b = (int)blockIdx.x - j * nbmax;
    const SellJob q = tab[j];
    if (b >= q.nb) return;
    sell_job_block<V>(*q.A, q, b);
}

// ------------------------------------------------------- split kernel --
// Implicit bases with a particle number (no symmetry, or spin inversion at
// half filling) and two-bit hop couplings.  A state splits into its high
// word h = s >> H and low word l = s & (2^H - 1).  In colex order the rows
// sharing h form one contiguous block of C(H, p) rows (p = nup - popcount h)
// ordered by the rank of l among the p-bit words, starting at bstart:
//  * a hop inside h moves whole blocks: row r of block h pulls row r of
//    block h ^ flip (row C(H,p)-1-r of the inverted block when the hop sets
//    bit L-1 under spin inversion).  Liveness, weight and the neighbour
//    block's offset are wavefront-uniform (scalar registers) and the reads
//    are contiguous;
//  * a hop inside l stays in the block, staged in LDS: the neighbour's row
//    is rank[l ^ flip] from an LDS table over every H-bit word;
//  * a hop across the split changes p and reaches one whole block at rows
//    rank[l'] in any order: that block is staged in LDS beside the own one
//    (up to the LDS budget; further ones are gathered from HBM).
// Each wavefront owns one block at a time (at most 252 rows, four per lane)
// and needs no workgroup barrier, so every wavefront of a CU has its own
// block's loads in flight.  The grid sweeps the blocks in ascending order:
// blocks a hop reaches behind the sweep front were read recently and sit in
// the shared MALL.
constexpr int kSpBlock = 256;
constexpr int kSpRows = 4;       // rows per lane: blocks of at most 256 rows
constexpr int kSpMaxHops = 64;   // high + cross hop groups, one lane each

struct SplitDev {
    int H;
    uint32_t lmask, wmask;
    const uint32_t *blist;       // high word of each block, ascending
    const uint32_t *bstart;      // first row of each block
    int nblocks;
    const uint32_t *boff;        // first row of the block of each high word (2^W)
    const uint16_t *rank;
    const uint16_t *words;
    int cls_off[34], cls_n[34];
    int nlow, nhigh, ncross;
    int chunk;                   // blocks per XCD chunk
    int own;                     // the block's own terms: diagonal, low hops, user diagonal
    int nstage;                  // cross blocks staged in LDS per wavefront
    int xstride;                 // elements per staged block
    int wave_lds;                // bytes of one wavefront's LDS region
    int off_rank, off_words, off_wave, off_xw;   // LDS layout (bytes)
};

// one live high / cross hop of the current block (wavefront-private LDS)
struct SpHop {
    uint32_t off;      // first row of the neighbour block
    uint32_t flags;    // bit 0: reversed rows (high) / inverted low word (cross); bit 1: h's bit hi
                       // (cross); bits 8..15: the low bit (cross); bits 16..31: rows of the block
    double re, im;     // weight: conjugated amplitude with the inversion character
};

template <typename V>
__device__ __forceinline__ void acc_nb(double wr, double wi, double vr, double vi, double &ar, double &ai) {
    if (VTraits<V>::cplx) {
        ar += wr * vr - wi * vi;
        ai += wr * vi + wi * vr;
    } else {
        ar += wr * vr;
    }
}

// Row-sum type of the split kernels: FP32 storage sums in FP32 (the FP32
// stream's products; its dots, norms and Rayleigh-Ritz step stay FP64),
// FP64 storage in FP64.
template <typename V> struct SplitAcc { using type = double; };
template <> struct SplitAcc<float> { using type = float; };
template <> struct SplitAcc<cplx_f> { using type = float; };

template <typename V, typename R>
__device__ __forceinline__ void acc_r(R wr, R wi, R vr, R vi, R &ar, R &ai) {
    if (VTraits<V>::cplx) {
        ar += wr * vr - wi * vi;
        ai += wr * vi + wi * vr;
    } else {
        ar += wr * vr;
    }
}

// LDS written by some lanes of a wavefront, read by others
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename V, bool CC, int CM>
__global__ __launch_bounds__(kSpBlock) void k_matvec_split(MVArgs A, SplitDev D, const HopDev *__restrict__ hops,
                                                           const DiagClass *__restrict__ dcls,
                                                           const TermDev *__restrict__ drest,
                                                           const V *__restrict__ x, V *__restrict__ y) {
    using T = VTraits<V>;
    // high hops per batch (x kSpRows loads in flight)
    constexpr int kNB = sizeof(V) <= 4 ? 10 : (sizeof(V) <= 8 ? 4 : 2);
    extern __shared__ __align__(16) char smem[];
    uint16_t *s_rank = (uint16_t *)(smem + D.off_rank);
    uint16_t *s_words = (uint16_t *)(smem + D.off_words);
    const int nw = 1 << D.H;
    for (int i = threadIdx.x; i < nw; i += kSpBlock) {
        s_rank[i] = D.rank[i];
        s_words[i] = D.words[i];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    SpHop *s_list = (SpHop *)(smem + D.off_wave + wave * D.wave_lds);
    V *s_x = (V *)(smem + D.off_wave + wave * D.wave_lds + D.off_xw);
    // lane k < nhigh owns high hop group k, lane nhigh + k cross group k,
    // for every block: their constants stay in registers
    const bool is_high = lane < D.nhigh;
    const bool is_cross = !is_high && lane < D.nhigh + D.ncross;
    int g_lo = 0, g_hi = 0, g_nz = 0, g_sym = 0, g_top = 0;
    double g_re1 = 0, g_re2 = 0, g_im1 = 0, g_im2 = 0;
    if (is_high || is_cross) {
        const HopDev &hg = hops[D.nlow + lane];
        g_lo = hg.lo;
        g_hi = hg.hi;
        g_nz = hg.nzmask;
        g_sym = hg.sym;
        g_top = hg.top;
        g_re1 = hg.re[1];
        g_re2 = hg.re[2];
        g_im1 = hg.im[1];
        g_im2 = hg.im[2];
    }
    double dot_r = 0, dot_i = 0;
    // Blocks come in chunks of D.chunk; chunk c belongs to XCD c % 8 (the
    // grid is a multiple of 8 and hardware places workgroup b on XCD b % 8),
    // and the wavefronts of one XCD deal its chunks out in order.  The eight
    // XCDs advance side by side, so the chip sweeps the blocks in ascending
    // order (far hops find their blocks in the MALL) while a block's near
    // neighbours were read on the same XCD (its L2).
    const int xcd = blockIdx.x & 7;
    const int nwx = (int)(gridDim.x >> 3) * (kSpBlock / 64);
    const int64_t C = D.chunk;
    auto block_of = [&](int64_t q) { return (8 * (q / C) + xcd) * C + q % C; };
    int64_t q = (int64_t)(blockIdx.x >> 3) * (kSpBlock / 64) + wave;
    int64_t vb = block_of(q);
    uint32_t hn = 0, on = 0;
    if (vb < D.nblocks) {
        hn = D.blist[vb];
        on = D.bstart[vb];
    }
    while (vb < D.nblocks) {
        const uint32_t h = hn, off = on;
        q += nwx;
        vb = block_of(q);
        if (vb < D.nblocks) {   // the next block's header, in flight meanwhile
            hn = D.blist[vb];
            on = D.bstart[vb];
        }
        if (h == 0xFFFFFFFFu) continue;   // padding of a tile list
        const int p = A.nup - __popc(h);
        const int n = D.cls_n[p];
        const int cw = D.cls_off[p];
        // own block (only the pass that applies the block's own terms reads it)
        double xo_r[kSpRows], xo_i[kSpRows];
#pragma unroll
        for (int j = 0; j < kSpRows; j++) {
            const int r = lane + 64 * j;
            xo_r[j] = 0;
            xo_i[j] = 0;
            if (D.own && r < n) T::load(x, (int64_t)off + r, xo_r[j], xo_i[j]);
            if (A.conj_x) xo_i[j] = -xo_i[j];
        }
        // this block's entry of the lane's hop group
        bool live = false;
        SpHop e{};
        if (is_high) {
            const int lo = g_lo - D.H, hi = g_hi - D.H;
            const uint32_t blo = (h >> lo) & 1, bhi = (h >> hi) & 1;
            live = (g_nz >> (blo | (bhi << 1))) & 1;
            // with a particle number only patterns 1 (bit lo set) and 2 are live
            double m_r = (g_sym || blo) ? g_re1 : g_re2;
            double m_i = CC ? ((g_sym || blo) ? g_im1 : g_im2) : 0.0;
            uint32_t h2 = h ^ ((1u << lo) | (1u << hi));
            uint32_t fl = 0;
            if (CM == CM_ZONLY && g_top) {
                // the hop sets bit L-1: the column is the inverted state's
                h2 = ~h2 & D.wmask;
                fl = 1;
                if (A.zchar < 0) {
                    m_r = -m_r;
                    m_i = -m_i;
                }
            }
            e.off = live ? D.boff[h2] : 0;
            e.flags = fl;
            e.re = m_r;
            e.im = -m_i;   // conj(amp): the terms are the adjoint couplings
        } else if (is_cross) {
            // the rows whose bit lo differs from h's bit hi move
            const int hi = g_hi - D.H;
            const uint32_t bhi = (h >> hi) & 1;
            const int pat = bhi ? 2 : 1;
            double m_r = bhi ? g_re2 : g_re1;
            double m_i = CC ? (bhi ? g_im2 : g_im1) : 0.0;
            uint32_t h2 = h ^ (1u << hi);
            const int inv = CM == CM_ZONLY && g_top;
            if (inv) {
                h2 = ~h2 & D.wmask;
                if (A.zchar < 0) {
                    m_r = -m_r;
                    m_i = -m_i;
                }
            }
            const int p2 = A.nup - __popc(h2);
            live = ((g_nz >> pat) & 1) && p2 >= 0 && p2 <= D.H;
            const uint32_t o2 = live ? D.boff[h2] : 0xFFFFFFFFu;
            live = live && o2 != 0xFFFFFFFFu;   // else no row of this block moves that way
            e.off = o2;
            e.flags = (uint32_t)inv | (bhi << 1) | ((uint32_t)g_lo << 8) |
                      ((uint32_t)(live ? D.cls_n[p2] : 0) << 16);
            e.re = m_r;
            e.im = -m_i;
        }
        const uint64_t mh = __ballot(is_high && live), mc = __ballot(is_cross && live);
        const uint64_t below = (1ULL << lane) - 1;
        if (is_high && live) s_list[__popcll(mh & below)] = e;
        if (is_cross && live) s_list[D.nhigh + __popcll(mc & below)] = e;
        const int nh = __popcll(mh), nc = __popcll(mc);
        if (D.own) {
#pragma unroll
            for (int j = 0; j < kSpRows; j++) {
                const int r = lane + 64 * j;
                if (r < D.xstride) T::store(s_x, r, xo_r[j], xo_i[j]);   // s_x[n] = 0: dead neighbours read it
            }
        }
        wave_lds_sync();
        // One round of loads for the first kNB high hops and the staged cross
        // blocks together; later batches of high hops (if any) follow.
        double ar[kSpRows], ai[kSpRows];
#pragma unroll
        for (int j = 0; j < kSpRows; j++) {
            ar[j] = 0;
            ai[j] = 0;
        }
        const int ns = min(nc, D.nstage);
        auto stage_cross = [&]() {
            for (int c = 0; c < ns; c++) {
                const SpHop ec = s_list[D.nhigh + c];
                const uint32_t n2 = ec.flags >> 16;
                V u[kSpRows];
#pragma unroll
                for (int j = 0; j < kSpRows; j++) {
                    const int i = lane + 64 * j;
                    u[j] = V{};
                    if (i < (int)n2) u[j] = x[(int64_t)ec.off + i];
                }
#pragma unroll
                for (int j = 0; j < kSpRows; j++) {
                    const int i = lane + 64 * j;
                    if (i < D.xstride) {
                        double ur, ui;
                        T::load(&u[j], 0, ur, ui);
                        T::store(s_x, (int64_t)(1 + c) * D.xstride + i, ur, A.conj_x ? -ui : ui);
                    }
                }
            }
        };
        for (int k0 = 0; k0 == 0 || k0 < nh; k0 += kNB) {
            if constexpr (sizeof(V) > 4) {
                // 8- and 16-byte elements: widened to FP64 as they land
                double vr[kNB][kSpRows], vi[kNB][kSpRows];
#pragma unroll
                for (int kk = 0; kk < kNB; kk++) {
                    const int k = max(0, min(k0 + kk, nh - 1));
                    const uint32_t eoff = __builtin_amdgcn_readfirstlane(s_list[k].off);
                    const uint32_t rev = __builtin_amdgcn_readfirstlane(s_list[k].flags) & 1;
                    const V *xb = x + eoff;
#pragma unroll
                    for (int j = 0; j < kSpRows; j++) {
                        const int r = lane + 64 * j;
                        vr[kk][j] = 0;
                        vi[kk][j] = 0;
                        if (k0 + kk < nh && r < n) T::load(xb, rev ? n - 1 - r : r, vr[kk][j], vi[kk][j]);
                    }
                }
                if (k0 == 0) stage_cross();
#pragma unroll
                for (int kk = 0; kk < kNB; kk++) {
                    if (k0 + kk < nh) {
                        const double wr = s_list[k0 + kk].re, wi = s_list[k0 + kk].im;
#pragma unroll
                        for (int j = 0; j < kSpRows; j++)
                            acc_nb<V>(wr, wi, vr[kk][j], A.conj_x ? -vi[kk][j] : vi[kk][j], ar[j], ai[j]);
                    }
                }
            } else {
                // 4-byte elements: raw values, one register each
                V raw[kNB][kSpRows];
#pragma unroll
                for (int kk = 0; kk < kNB; kk++) {
                    const int k = max(0, min(k0 + kk, nh - 1));
                    const uint32_t eoff = __builtin_amdgcn_readfirstlane(s_list[k].off);
                    const uint32_t rev = __builtin_amdgcn_readfirstlane(s_list[k].flags) & 1;
                    const V *xb = x + eoff;
#pragma unroll
                    for (int j = 0; j < kSpRows; j++) {
                        const int r = lane + 64 * j;
                        if (k0 + kk < nh && r < n) raw[kk][j] = xb[rev ? n - 1 - r : r];
                    }
                }
                if (k0 == 0) stage_cross();
#pragma unroll
                for (int kk = 0; kk < kNB; kk++) {
                    if (k0 + kk < nh) {
                        const double wr = s_list[k0 + kk].re, wi = s_list[k0 + kk].im;
#pragma unroll
                        for (int j = 0; j < kSpRows; j++) {
                            if (lane + 64 * j < n) {
                                double vr, vi;
                                T::load(&raw[kk][j], 0, vr, vi);
                                acc_nb<V>(wr, wi, vr, A.conj_x ? -vi : vi, ar[j], ai[j]);
                            }
                        }
                    }
                }
            }
        }
        wave_lds_sync();
        // every row: diagonal, low hops, cross hops, in that order; the loops
        // run group by group over the lane's rows, so each group's constants
        // are loaded once per block
        uint32_t lw[kSpRows];
#pragma unroll
        for (int j = 0; j < kSpRows; j++) {
            const int r = lane + 64 * j;
            lw[j] = 0;
            if (r >= n) continue;
            const uint32_t l = s_words[cw + r];
            lw[j] = l;
            if (!D.own) continue;
            double dr = 0, di = 0;
            diag_value<CC>(A, dcls, drest, (h << D.H) | l, dr, di);
            di = -di;   // conj(amp): the terms are the adjoint couplings
            if (A.user_diag) {
                const int64_t row = (int64_t)off + r;
                dr += A.user_diag[2 * row];
                di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
            }
            acc_nb<V>(dr, di, xo_r[j], xo_i[j], ar[j], ai[j]);
        }
        // low hops: neighbours in the staged own block
        for (int g = 0; g < D.nlow; g++) {
            const HopDev hp = hops[g];
#pragma unroll
            for (int j = 0; j < kSpRows; j++) {
                if (lane + 64 * j >= n) continue;
                const uint32_t l = lw[j];
                const uint32_t blo = (l >> hp.lo) & 1, bhi = (l >> hp.hi) & 1;
                const bool lv = (hp.nzmask >> (blo | (bhi << 1))) & 1;
                // with a particle number only patterns 1 and 2 are live
                const double m_r = hp.sym ? hp.re[1] : (blo ? hp.re[1] : hp.re[2]);
                const double m_i = CC ? (hp.sym ? hp.im[1] : (blo ? hp.im[1] : hp.im[2])) : 0.0;
                const int nid = lv ? (int)s_rank[l ^ ((1u << hp.lo) | (1u << hp.hi))] : n;
                double vr, vi;
                T::load(s_x, nid, vr, vi);
                acc_nb<V>(m_r, -m_i, vr, vi, ar[j], ai[j]);
            }
        }
        // cross hops: staged neighbour blocks (or HBM gathers past the budget)
        for (int c = 0; c < nc; c++) {
            const SpHop ec = s_list[D.nhigh + c];
            const uint32_t lo = (ec.flags >> 8) & 0xFF, bhi = (ec.flags >> 1) & 1;
#pragma unroll
            for (int j = 0; j < kSpRows; j++) {
                if (lane + 64 * j >= n) continue;
                const uint32_t l = lw[j];
                if (((l >> lo) & 1) == bhi) continue;
                uint32_t l2 = l ^ (1u << lo);
                if (ec.flags & 1) l2 = ~l2 & D.lmask;
                const int idx = s_rank[l2];
                double vr, vi;
                if (c < D.nstage) {
                    T::load(s_x, (int64_t)(1 + c) * D.xstride + idx, vr, vi);
                } else {
                    T::load(x, (int64_t)ec.off + idx, vr, vi);
                    if (A.conj_x) vi = -vi;
                }
                acc_nb<V>(ec.re, ec.im, vr, vi, ar[j], ai[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < kSpRows; j++) {
            const int r = lane + 64 * j;
            if (r < n) finish_row<V>(A, y, (int64_t)off + r, ar[j], ai[j], 0.0, 0.0, dot_r, dot_i);
        }
        wave_lds_sync();   // the block's LDS reads are done before the next block's writes
    }
}

// ------------------------------------------------- lean split kernel --
// The same block decomposition with the per-row bookkeeping moved into
// tables and scalar registers:
//  * the low hops' neighbour rows come from a per-basis, per-coupling-list
//    LDS table indexed by the row's position in its popcount class (one byte
//    per hop, 255 = no neighbour: row 255 of the staged block is always 0);
//  * the cross hops' rows in the neighbour block come from a second table
//    (the row's bit lo in bit 15), liveness from that bit and h's bit hi;
//  * the high hops are found by one ballot over lanes that hold one hop each;
//    each live hop's neighbour-block offset and weight move to scalar
//    registers (readlane), so a neighbour row is one load with a scalar base
//    and the lane's fixed byte offset, and one FMA;
//  * row slots past the block's last row are skipped by wave-uniform branches.
// All high-hop loads of a batch are issued before the own block's diagonal
// and low hops are computed, so their latency overlaps that arithmetic.
constexpr int kS2MaxLow = 32;    // low hop groups
constexpr int kS2MaxCross = 16;  // cross hop groups

struct Split2Dev {
    int H;
    uint32_t lmask, wmask;
    const uint32_t *blist, *bstart;
    int nblocks, chunk, own;
    const uint32_t *boff;
    int cls_off[34], cls_n[34];
    int nlow, nhigh, ncross;
    int sl, sc;              // table strides: nlow rounded up to 4, ncross rounded up to 2
    const uint16_t *words;   // H-bit words by (popcount, value)
    const uint8_t *lnb;      // [w * sl + g]: low hop g's neighbour row of class position w (255: none)
    const uint16_t *cnb;     // [w * sc + c]: cross hop c's row in the neighbour block | bit lo << 15
    int off_words, off_lnb, off_cnb, off_wave, wave_bytes;
};

// Per-lane data of the lean kernel: the high hop group this lane holds.
struct S2Lane {
    int lo, hi, nz, sym, top;
    double re1, re2, im1, im2;
};

// One block of the lean kernel with NS row slots of 64 rows (n <= 64 NS).
// Every load is issued unconditionally: rows past the block's end read its
// last row and hops past a batch's end read the own block (weight 0), so no
// load sits in a branch that would make the compiler wait for it there.
template <typename V, bool CC, int CM, int NS>
__device__ __forceinline__ void split2_block(const MVArgs &A, const Split2Dev &D, const HopDev *__restrict__ hops,
                                             const DiagClass *__restrict__ dcls, const TermDev *__restrict__ drest,
                                             const V *__restrict__ x, V *__restrict__ y, const uint16_t *s_words,
                                             const uint8_t *s_lnb, const uint16_t *s_cnb, V *s_x, const S2Lane &g,
                                             bool is_high, int lane, uint32_t h, uint32_t off, int n, int cw) {
    using T = VTraits<V>;
    using R = typename SplitAcc<V>::type;   // row sums: FP32 for FP32 storage
    constexpr int kNB = sizeof(V) <= 4 ? 6 : (sizeof(V) <= 8 ? 5 : 3);
    constexpr uint32_t ES = sizeof(V);
    // this block's high hops: liveness by ballot; neighbour offset, row
    // order and weight stay in the lane that holds the hop
    bool live = false;
    uint32_t e_off = 0;
    double e_re = 0, e_im = 0;
    int e_rev = 0;
    if (is_high) {
        const uint32_t blo = (h >> g.lo) & 1, bhi = (h >> g.hi) & 1;
        live = (g.nz >> (blo | (bhi << 1))) & 1;
        e_re = (g.sym || blo) ? g.re1 : g.re2;
        e_im = CC ? -((g.sym || blo) ? g.im1 : g.im2) : 0.0;   // conj(amp): adjoint couplings
        uint32_t h2 = h ^ ((1u << g.lo) | (1u << g.hi));
        if (CM == CM_ZONLY && g.top) {
            // the hop sets bit L-1: the inverted state's column, rows reversed
            h2 = ~h2 & D.wmask;
            e_rev = 1;
            if (A.zchar < 0) {
                e_re = -e_re;
                e_im = -e_im;
            }
        }
        e_off = live ? D.boff[h2] : off;
    }
    // hops in row order and hops onto the inverted block (rows reversed)
    uint64_t lm = __ballot(live && !e_rev);
    uint64_t lmr = __ballot(live && e_rev);
    uint32_t roff[NS], rrev[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const int r = min(lane + 64 * j, n - 1);
        roff[j] = (uint32_t)r * ES;
        rrev[j] = (uint32_t)(n - 1 - r) * ES;
    }
    const V *xo_base = x + off;
    V xo[NS];
    if (D.own) {
#pragma unroll
        for (int j = 0; j < NS; j++) xo[j] = *atb(xo_base, roff[j]);
    }
    R ar[NS], ai[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) {
        ar[j] = 0;
        ai[j] = 0;
    }
    bool first = true;
    do {
        // one batch of up to kNB live high hops (lane indices in scalars)
        int k_idx[kNB];
        int nb = 0;
#pragma unroll
        for (int kk = 0; kk < kNB; kk++) {
            k_idx[kk] = 0;
            if (lm) {
                k_idx[kk] = __builtin_ctzll(lm);
                lm &= lm - 1;
                nb = kk + 1;
            }
        }
        V v[kNB][NS];
#pragma unroll
        for (int kk = 0; kk < kNB; kk++) {
            const uint32_t bo = __builtin_amdgcn_readlane(e_off, k_idx[kk]);
            const V *base = x + (kk < nb ? bo : off);
#pragma unroll
            for (int j = 0; j < NS; j++) v[kk][j] = *atb(base, roff[j]);
        }
        if (first) {
            first = false;
            if (D.own) {
                // own block: stage in LDS, diagonal, low hops (the high-hop
                // loads above are in flight meanwhile)
#pragma unroll
                for (int j = 0; j < NS; j++)
                    if (lane + 64 * j < n) s_x[lane + 64 * j] = xo[j];
                wave_lds_sync();
                uint32_t lw[NS];
#pragma unroll
                for (int j = 0; j < NS; j++) {
                    const int r = min(lane + 64 * j, n - 1);
                    double xr, xi;
                    T::load(&xo[j], 0, xr, xi);
                    if (A.conj_x) xi = -xi;
                    const uint32_t l = s_words[cw + r];
                    lw[j] = l;
                    double dr = 0, di = 0;
                    diag_value<CC>(A, dcls, drest, (h << D.H) | l, dr, di);
                    di = -di;   // conj(amp): the terms are the adjoint couplings
                    if (A.user_diag) {
                        const int64_t row = (int64_t)off + r;
                        dr += A.user_diag[2 * row];
                        di += A.user_diag_conj ? -A.user_diag[2 * row + 1] : A.user_diag[2 * row + 1];
                    }
                    acc_r<V, R>((R)dr, (R)di, (R)xr, (R)xi, ar[j], ai[j]);
                }
                for (int gi = 0; gi < D.nlow; gi++) {
                    const HopDev &hp = hops[gi];
                    const R r1 = (R)hp.re[1], r2 = (R)hp.re[2];
                    const R i1 = CC ? (R)hp.im[1] : (R)0, i2 = CC ? (R)hp.im[2] : (R)0;
                    const int sym = hp.sym, lo = hp.lo;
#pragma unroll
                    for (int j = 0; j < NS; j++) {
                        const int r = min(lane + 64 * j, n - 1);
                        const uint32_t nid = s_lnb[(cw + r) * D.sl + gi];
                        double vr, vi;
                        T::load(s_x, nid, vr, vi);
                        if (A.conj_x) vi = -vi;
                        R mr = r1, mi = i1;
                        if (!sym && !((lw[j] >> lo) & 1)) {
                            mr = r2;
                            mi = i2;
                        }
                        acc_r<V, R>(mr, -mi, (R)vr, (R)vi, ar[j], ai[j]);
                    }
                }
            }
        }
#pragma unroll
        for (int kk = 0; kk < kNB; kk++) {
            if (kk < nb) {
                const uint64_t rb = __builtin_bit_cast(uint64_t, e_re);
                const double wr = __builtin_bit_cast(
                    double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(rb >> 32), k_idx[kk]) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((uint32_t)rb, k_idx[kk]));
                double wi = 0;
                if (CC) {
                    const uint64_t ib = __builtin_bit_cast(uint64_t, e_im);
                    wi = __builtin_bit_cast(
                        double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(ib >> 32), k_idx[kk]) << 32) |
                                    (uint32_t)__builtin_amdgcn_readlane((uint32_t)ib, k_idx[kk]));
                }
#pragma unroll
                for (int j = 0; j < NS; j++) {
                    double vr, vi;
                    T::load(&v[kk][j], 0, vr, vi);
                    acc_r<V, R>((R)wr, (R)wi, (R)vr, (R)(A.conj_x ? -vi : vi), ar[j], ai[j]);
                }
            }
        }
    } while (lm);
    // hops that set bit L-1 under spin inversion: row r of the inverted block
    // is row n-1-r
    while (lmr) {
        const int k = __builtin_ctzll(lmr);
        lmr &= lmr - 1;
        const V *base = x + __builtin_amdgcn_readlane(e_off, k);
        V v[NS];
#pragma unroll
        for (int j = 0; j < NS; j++) v[j] = *atb(base, rrev[j]);
        const uint64_t rb = __builtin_bit_cast(uint64_t, e_re);
        const double wr = __builtin_bit_cast(
            double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(rb >> 32), k) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((uint32_t)rb, k));
        double wi = 0;
        if (CC) {
            const uint64_t ib = __builtin_bit_cast(uint64_t, e_im);
            wi = __builtin_bit_cast(
                double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(ib >> 32), k) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((uint32_t)ib, k));
        }
#pragma unroll
        for (int j = 0; j < NS; j++) {
            double vr, vi;
            T::load(&v[j], 0, vr, vi);
            acc_r<V, R>((R)wr, (R)wi, (R)vr, (R)(A.conj_x ? -vi : vi), ar[j], ai[j]);
        }
    }
    // cross hops: rows of the neighbour block from the table, gathered
    for (int c = 0; c < D.ncross; c++) {
        const HopDev &hc = hops[D.nlow + D.nhigh + c];
        const int hi = hc.hi - D.H;
        const uint32_t bhi = (h >> hi) & 1;
        double m_r = bhi ? hc.re[2] : hc.re[1];
        double m_i = CC ? (bhi ? hc.im[2] : hc.im[1]) : 0.0;
        uint32_t h2 = h ^ (1u << hi);
        if (CM == CM_ZONLY && hc.top) {
            h2 = ~h2 & D.wmask;
            if (A.zchar < 0) {
                m_r = -m_r;
                m_i = -m_i;
            }
        }
        const int p2 = A.nup - __popc(h2);
        if (!((hc.nzmask >> (bhi ? 2 : 1)) & 1) || p2 < 0 || p2 > D.H) continue;
        const uint32_t o2 = D.boff[h2];
        if (o2 == 0xFFFFFFFFu) continue;
        const V *b2 = x + o2;
        V u[NS];
        bool lv[NS];
#pragma unroll
        for (int j = 0; j < NS; j++) {
            const int r = min(lane + 64 * j, n - 1);
            const uint32_t e = s_cnb[(cw + r) * D.sc + c];
            lv[j] = (e >> 15) != bhi;
            u[j] = *atb(b2, (lv[j] ? (e & 0x7FFFu) : 0u) * ES);
        }
#pragma unroll
        for (int j = 0; j < NS; j++) {
            double vr, vi;
            T::load(&u[j], 0, vr, vi);
            if (!lv[j]) {
                vr = 0;
                vi = 0;
            }
            acc_r<V, R>((R)m_r, (R)(-m_i), (R)vr, (R)(A.conj_x ? -vi : vi), ar[j], ai[j]);
        }
    }
    double dot_r = 0, dot_i = 0;
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const int r = lane + 64 * j;
        if (r < n) finish_row<V>(A, y, (int64_t)off + r, (double)ar[j], (double)ai[j], 0.0, 0.0, dot_r, dot_i);
    }
}

template <typename V, bool CC, int CM>
__global__ __launch_bounds__(kSpBlock) void k_matvec_split2(MVArgs A, Split2Dev D, const HopDev *__restrict__ hops,
                                                            const DiagClass *__restrict__ dcls,
                                                            const TermDev *__restrict__ drest,
                                                            const V *__restrict__ x, V *__restrict__ y) {
    using T = VTraits<V>;
    extern __shared__ __align__(16) char smem[];
    uint16_t *s_words = (uint16_t *)(smem + D.off_words);
    uint8_t *s_lnb = (uint8_t *)(smem + D.off_lnb);
    uint16_t *s_cnb = (uint16_t *)(smem + D.off_cnb);
    const int nw = 1 << D.H;
    for (int i = threadIdx.x; i < nw; i += kSpBlock) s_words[i] = D.words[i];
    if (D.own && D.nlow)
        for (int i = threadIdx.x; i < nw * D.sl / 4; i += kSpBlock)
            ((uint32_t *)s_lnb)[i] = ((const uint32_t *)D.lnb)[i];
    if (D.ncross)
        for (int i = threadIdx.x; i < nw * D.sc / 2; i += kSpBlock)
            ((uint32_t *)s_cnb)[i] = ((const uint32_t *)D.cnb)[i];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    V *s_x = (V *)(smem + D.off_wave + wave * D.wave_bytes);
    if (lane == 0) T::store(s_x, 255, 0.0, 0.0);   // the row dead low-hop neighbours read
    __syncthreads();
    // lane k < nhigh holds high hop group k
    const bool is_high = lane < D.nhigh;
    S2Lane g{0, 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0};
    if (is_high) {
        const HopDev &hg = hops[D.nlow + lane];
        g.lo = hg.lo - D.H;
        g.hi = hg.hi - D.H;
        g.nz = hg.nzmask;
        g.sym = hg.sym;
        g.top = hg.top;
        g.re1 = hg.re[1];
        g.re2 = hg.re[2];
        g.im1 = hg.im[1];
        g.im2 = hg.im[2];
    }
    const int xcd = blockIdx.x & 7;
    const int nwx = (int)(gridDim.x >> 3) * (kSpBlock / 64);
    const int64_t C = D.chunk;
    auto block_of = [&](int64_t q) { return (8 * (q / C) + xcd) * C + q % C; };
    int64_t q = (int64_t)(blockIdx.x >> 3) * (kSpBlock / 64) + wave;
    int64_t vb = block_of(q);
    uint32_t hn = 0, on = 0;
    if (vb < D.nblocks) {
        hn = D.blist[vb];
        on = D.bstart[vb];
    }
    while (vb < D.nblocks) {
        const uint32_t h = __builtin_amdgcn_readfirstlane(hn), off = __builtin_amdgcn_readfirstlane(on);
        q += nwx;
        vb = block_of(q);
        if (vb < D.nblocks) {   // the next block's header, in flight meanwhile
            hn = D.blist[vb];
            on = D.bstart[vb];
        }
        if (h == 0xFFFFFFFFu) continue;   // padding of a tile list
        const int p = A.nup - __popc(h);
        const int n = D.cls_n[p];
        const int cw = D.cls_off[p];
        const int ns = (n + 63) >> 6;   // row slots in use (wave-uniform)
        if (ns <= 1)
            split2_block<V, CC, CM, 1>(A, D, hops, dcls, drest, x, y, s_words, s_lnb, s_cnb, s_x, g, is_high, lane, h,
                                       off, n, cw);
        else if (ns == 2)
            split2_block<V, CC, CM, 2>(A, D, hops, dcls, drest, x, y, s_words, s_lnb, s_cnb, s_x, g, is_high, lane, h,
                                       off, n, cw);
        else if (ns == 3)
            split2_block<V, CC, CM, 3>(A, D, hops, dcls, drest, x, y, s_words, s_lnb, s_cnb, s_x, g, is_high, lane, h,
                                       off, n, cw);
        else
            split2_block<V, CC, CM, 4>(A, D, hops, dcls, drest, x, y, s_words, s_lnb, s_cnb, s_x, g, is_high, lane, h,
                                       off, n, cw);
        wave_lds_sync();   // the block's LDS reads are done before the next block's writes
    }
}

struct Launch {
    int grid;
    size_t lds;
    hipStream_t s;
    const GroupOfTerms *grp;
    const TermDev *off, *diag;
};

struct HopLaunch {
    const HopDev *hops;
    const DiagClass *dcls;
    const TermDev *drest;
};

template <typename V, bool CC>
void dispatch_hop_cm(const MVArgs &a, int cm, bool track, int ws, const void *x, const void *xs, void *y,
                     const Launch &L, const HopLaunch &H) {
    using W = typename Wide<V>::type;
    const V *xv = (const V *)x;
    const W *xw = (const W *)xs;
    V *yv = (V *)y;
    // xs given (parent-table mode): neighbours gathered from the pre-scaled x
#define QS_HOP_WS(CMV, TR, P, W)                                                                                \
    {                                                                                                              \
        if constexpr (CMV == CM_PTAB) {                                                                            \
            if (xw) {                                                                                              \
                hipLaunchKernelGGL((k_matvec_hop<V, CC, CMV, TR, P, W, true>), dim3(L.grid), dim3(kMvBlock),       \
                                   L.lds, L.s, a, H.hops, H.dcls, H.drest, xv, xw, yv);                            \
                break;                                                                                             \
            }                                                                                                      \
        }                                                                                                          \
        hipLaunchKernelGGL((k_matvec_hop<V, CC, CMV, TR, P, W, false>), dim3(L.grid), dim3(kMvBlock), L.lds, L.s,  \
                           a, H.hops, H.dcls, H.drest, xv, nullptr, yv);                                           \
        break;                                                                                                     \
    }
#define QS_HOP(CMV, TR, P)                                                                                      \
    switch (ws) {                                                                                                  \
    case 1: QS_HOP_WS(CMV, TR, P, 1)                                                                               \
    case 2: QS_HOP_WS(CMV, TR, P, 2)                                                                               \
    default: QS_HOP_WS(CMV, TR, P, 4)                                                                              \
    }
    if (a.nup >= 0) {
        if (cm == CM_FULL) QS_HOP(CM_FULL, false, true)
        else if (cm == CM_ZONLY) QS_HOP(CM_ZONLY, false, true)
        else if (track) QS_HOP(CM_PTAB, true, true)
        else QS_HOP(CM_PTAB, false, true)
    } else {
        if (cm == CM_FULL) QS_HOP(CM_FULL, false, false)
        else if (cm == CM_ZONLY) QS_HOP(CM_ZONLY, false, false)
        else if (track) QS_HOP(CM_PTAB, true, false)
        else QS_HOP(CM_PTAB, false, false)
    }
#undef QS_HOP
#undef QS_HOP_WS
}

// ps in the parent-table mode: x * (1/sqrt(D)) is written once to an FP64
// buffer the operator keeps for the launching stream (kept between products:
// a stream-ordered allocation per product cost more host time than the
// product itself) and the hop kernel gathers from it.  The buffer belongs to
// the stream, so the pre-scale pass and the gathers of one product are
// ordered by that stream alone and products on other streams use their own.
template <typename V, bool CC>
void hop_cm(const MVArgs &a, int cm, bool track, int ws, bool ps, const void *x, void *y, const Launch &L,
            const HopLaunch &H, const OpHost &op, void *xs_given) {
    if (ps && cm == CM_PTAB) {
        using W = typename Wide<V>::type;
        const int64_t n = a.N;
        // sized for the widest element so one buffer serves every dtype (or
        // the caller's buffer: the fixed scratch of the no-allocation product)
        const size_t bytes = sizeof(cplx_d) * (size_t)std::max<int64_t>(n, 1);
        W *xs = xs_given ? (W *)xs_given : (W *)op.xs_for(L.s, bytes);
        const int g = (int)std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL((k_prescale<V>), dim3(std::max(g, 1)), dim3(256), 0, L.s, (const V *)x, a.inv_nrm, xs, n);
        dispatch_hop_cm<V, CC>(a, cm, track, ws, x, xs, y, L, H);
        return;
    }
    dispatch_hop_cm<V, CC>(a, cm, track, ws, x, nullptr, y, L, H);
}

void dispatch_hop(const MVArgs &a, int dtype, bool cc, int cm, bool track, int ws, bool ps, const void *x, void *y,
                  const Launch &L, const HopLaunch &H, const OpHost &op, void *xs_given) {
    switch (dtype) {
    case QS_F64: hop_cm<double, false>(a, cm, track, ws, ps, x, y, L, H, op, xs_given); break;
    case QS_F32: hop_cm<float, false>(a, cm, track, ws, ps, x, y, L, H, op, xs_given); break;
    case QS_C128:
        if (cc) hop_cm<cplx_d, true>(a, cm, track, ws, ps, x, y, L, H, op, xs_given);
        else hop_cm<cplx_d, false>(a, cm, track, ws, ps, x, y, L, H, op, xs_given);
        break;
    case QS_C64:
        if (cc) hop_cm<cplx_f, true>(a, cm, track, ws, ps, x, y, L, H, op, xs_given);
        else hop_cm<cplx_f, false>(a, cm, track, ws, ps, x, y, L, H, op, xs_given);
        break;
    default: fail(QS_EARG, "unknown dtype");
    }
}

template <typename S, typename V, bool CC, bool TRACK, bool TLDS>
void dispatch_lpr(const MVArgs &a, int lpr, const void *x, void *y, const Launch &L) {
    const V *xv = (const V *)x;
    V *yv = (V *)y;
#define QS_STORED(N)                                                                                            \
    hipLaunchKernelGGL((k_matvec_stored<S, V, CC, TRACK, N, TLDS>), dim3(L.grid), dim3(kStBlock), L.lds, L.s, a, \
                       L.grp, L.off, L.diag, xv, yv)
    switch (lpr) {
    case 1: QS_STORED(1); break;
    case 2: QS_STORED(2); break;
    case 4: QS_STORED(4); break;
    default: QS_STORED(8); break;
    }
#undef QS_STORED
}

template <typename S, typename V, bool CC>
void dispatch_mode(const MVArgs &a, int mode, bool track, bool tlds, int lpr, const void *x, void *y,
                   const Launch &L) {
    const V *xv = (const V *)x;
    V *yv = (V *)y;
    if (mode == MODE_FULL) {
        hipLaunchKernelGGL((k_matvec_implicit<S, V, CC, MODE_FULL>), dim3(L.grid), dim3(kMvBlock), L.lds, L.s, a,
                           L.grp, L.off, L.diag, xv, yv);
    } else if (mode == MODE_ZONLY) {
        hipLaunchKernelGGL((k_matvec_implicit<S, V, CC, MODE_ZONLY>), dim3(L.grid), dim3(kMvBlock), L.lds, L.s, a,
                           L.grp, L.off, L.diag, xv, yv);
    } else if (track) {
        if (tlds) dispatch_lpr<S, V, CC, true, true>(a, lpr, x, y, L);
        else dispatch_lpr<S, V, CC, true, false>(a, lpr, x, y, L);
    } else {
        if (tlds) dispatch_lpr<S, V, CC, false, true>(a, lpr, x, y, L);
        else dispatch_lpr<S, V, CC, false, false>(a, lpr, x, y, L);
    }
}

template <typename S, typename V>
void dispatch_cc(const MVArgs &a, bool cc, int mode, bool track, bool tlds, int lpr, const void *x, void *y,
                 const Launch &L) {
    if constexpr (VTraits<V>::cplx) {
        if (cc) dispatch_mode<S, V, true>(a, mode, track, tlds, lpr, x, y, L);
        else dispatch_mode<S, V, false>(a, mode, track, tlds, lpr, x, y, L);
    } else {
        dispatch_mode<S, V, false>(a, mode, track, tlds, lpr, x, y, L);
    }
}

template <typename S>
void dispatch_v(const MVArgs &a, int dtype, bool cc, int mode, bool track, bool tlds, int lpr, const void *x,
                void *y, const Launch &L) {
    switch (dtype) {
    case QS_F64: dispatch_cc<S, double>(a, cc, mode, track, tlds, lpr, x, y, L); break;
    case QS_C128: dispatch_cc<S, cplx_d>(a, cc, mode, track, tlds, lpr, x, y, L); break;
    case QS_F32: dispatch_cc<S, float>(a, cc, mode, track, tlds, lpr, x, y, L); break;
    case QS_C64: dispatch_cc<S, cplx_f>(a, cc, mode, track, tlds, lpr, x, y, L); break;
    default: fail(QS_EARG, "unknown dtype");
    }
}

int grid_for(int64_t n, int rows_per_block) {
    int64_t blocks = (n + rows_per_block - 1) / rows_per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, (int64_t)1 << 22));
}

}  // namespace

int default_lanes_per_row(const BasisHost &b) { return b.mode == MODE_STORED ? 2 : 1; }

bool hop_eligible(const OpHost &op, bool adjoint) {
    const BasisHost &b = *op.basis;
    const FastTerms &f = adjoint ? op.fast_adj : op.fast;
    if (!f.hop_ok || b.use64 || b.Ns >= ((int64_t)1 << 31) || b.npar > ((int64_t)1 << 31)) return false;
    if (b.mode == MODE_STORED) return b.ptab.get() != nullptr;
    return true;
}

bool split_eligible(const OpHost &op, bool adjoint) {
    const BasisHost &b = *op.basis;
    const FastTerms &f = adjoint ? op.fast_adj : op.fast;
    return f.hop_ok && !b.use64 && b.nup >= 0 && b.L >= 8 && (b.mode == MODE_FULL || b.mode == MODE_ZONLY) &&
           b.Ns < ((int64_t)1 << 31);
}

// blocks of a matvec launch that carries FP64 dot partials (one per block):
// the stored kernel at LPR = 8, and the implicit kernel sized the same way
int matvec_grid(int64_t n) { return grid_for(n, kStBlock / 8); }

namespace {

MVArgs base_args(const OpHost &op, bool adjoint, double are, double aim, double bre, double bim, bool conj_x,
                 bool conj_y, double *dot_partial) {
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
    a.hkeys = b.hkeys.get();
    a.hvals = b.hvals.get();
    a.hmask = b.hmask;
    a.hmax_probe = b.hmax_probe;
    a.binom = b.binom.get();
    a.gchi = b.group.general ? b.gchi.get() : nullptr;
    // plain products use the adjoint couplings (pull form), adjoint products the couplings
    const std::vector<CTerm> &dg = adjoint ? op.diag_terms_adj : op.diag_terms;
    const std::vector<CTerm> &of = adjoint ? op.off_terms_adj : op.off_terms;
    const std::vector<GroupOfTerms> &gr = adjoint ? op.groups_adj : op.groups;
    a.ndiag = (int)dg.size();
    a.noff = (int)of.size();
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
    return a;
}

bool basis_tracks(const BasisHost &b) {
    if (b.group.general) {
        for (const cd &c : b.group.echar)
            if (c != cd(1, 0)) return true;
        return false;
    }
    for (int p = 0; p < b.group.npoint; p++)
        if (b.group.pchar[p] != cd(1, 0)) return true;
    for (int t = 0; t < b.group.nT; t++)
        if (b.group.lam[t] != cd(1, 0)) return true;
    return false;
}

int hop_mode(const BasisHost &b) {
    return b.mode == MODE_STORED ? CM_PTAB : (b.mode == MODE_ZONLY ? CM_ZONLY : CM_FULL);
}

// hop-kernel fields and LDS layout; returns the dynamic LDS bytes
size_t hop_args(MVArgs &a, const OpHost &op, bool adjoint, bool dot_partial) {
    const BasisHost &b = *op.basis;
    const FastTerms &ft = adjoint ? op.fast_adj : op.fast;
    a.reps = b.mode == MODE_STORED || b.states_stored ? (const void *)b.reps32.get() : nullptr;
    a.ptab = b.ptab.get();
    a.colbits = b.ptab_colbits;
    a.cfull = (b.nup >= 0) ? (uint32_t)binom(b.L, b.nup) : 0u;
    a.ndcls = (int)ft.dcls.size();
    a.ndrest = (int)ft.drest.size();
    size_t off = 0;
    auto place = [&](size_t bytes) {
        size_t o = (off + 15) & ~(size_t)15;
        off = o + bytes;
        return (int)o;
    };
    a.off_binom = place(b.nup >= 0 ? sizeof(uint32_t) * 33 * 65 : 0);
    a.off_red = place(dot_partial ? sizeof(double) * 2 * kMvBlock : 0);
    return off;
}

// split-kernel tables of the basis (once per basis)
const SplitBasis &split_basis(const BasisHost &b) {
    std::call_once(b.split_once, [&] {
        auto t = std::make_unique<SplitBasis>();
        const int L = b.L, nup = b.nup;
        const int H = L >= 20 ? 10 : std::max(2, L / 2);
        const int W = L - H;
        t->H = H;
        t->W = W;
        std::vector<uint16_t> words, rank((size_t)1 << H);
        for (int p = 0; p <= H; p++) {
            t->cls_off[p] = (int)words.size();
            for (uint32_t w = 0; w < (1u << H); w++)
                if (__builtin_popcount(w) == p) {
                    rank[w] = (uint16_t)(words.size() - t->cls_off[p]);
                    words.push_back((uint16_t)w);
                }
            t->cls_n[p] = (int)words.size() - t->cls_off[p];
            t->maxblock = std::max(t->maxblock, t->cls_n[p]);
        }
        std::vector<uint32_t> boff((size_t)1 << W, 0xFFFFFFFFu), blist, bstart;
        int64_t next = 0;
        for (uint32_t h = 0; h < (1u << W); h++) {
            const int p = nup - __builtin_popcount(h);
            if (p < 0 || p > H) continue;
            if (b.mode == MODE_ZONLY && ((h >> (W - 1)) & 1)) continue;
            // colex rank of the block's first state: h's particles sit above
            // the p lowest ones
            int64_t off = 0;
            int k = p;
            for (int i = 0; i < W; i++)
                if ((h >> i) & 1) off += binom(H + i, ++k);
            if (off != next) fail(QS_EARG, "internal: split blocks are not contiguous");
            boff[h] = (uint32_t)off;
            blist.push_back(h);
            bstart.push_back((uint32_t)off);
            next += t->cls_n[p];
        }
        if (next != b.Ns) fail(QS_EARG, "internal: split blocks do not cover the basis");
        t->nblocks = (int)blist.size();
        t->blist.alloc(blist.size());
        t->bstart.alloc(bstart.size());
        QS_HIP(hipMemcpy(t->bstart.get(), bstart.data(), bstart.size() * 4, hipMemcpyHostToDevice));
        t->boff.alloc(boff.size());
        t->rank.alloc(rank.size());
        t->words.alloc(words.size());
        QS_HIP(hipMemcpy(t->blist.get(), blist.data(), blist.size() * 4, hipMemcpyHostToDevice));
        QS_HIP(hipMemcpy(t->boff.get(), boff.data(), boff.size() * 4, hipMemcpyHostToDevice));
        QS_HIP(hipMemcpy(t->rank.get(), rank.data(), rank.size() * 2, hipMemcpyHostToDevice));
        QS_HIP(hipMemcpy(t->words.get(), words.data(), words.size() * 2, hipMemcpyHostToDevice));
        t->rank_h = std::move(rank);
        t->words_h = std::move(words);
        b.split = std::move(t);
    });
    return *b.split;
}

// The lean kernel's tables for hop list `o` (low groups first, cross
// groups last) on basis tables t.
void lean_tables(SplitHops &o, const std::vector<HopDev> &low, const std::vector<HopDev> &cross,
                 const SplitBasis &t, bool zonly) {
    o.lean = o.nlow <= kS2MaxLow && o.ncross <= kS2MaxCross && o.nhigh <= 64 && t.maxblock <= 255;
    if (!o.lean) return;
    const int nw = 1 << t.H;
    const uint32_t lmask = (1u << t.H) - 1;
    o.sl = std::max(4, (o.nlow + 3) & ~3);
    o.sc = std::max(2, (o.ncross + 1) & ~1);
    std::vector<uint8_t> lnb((size_t)nw * o.sl, 255);
    std::vector<uint16_t> cnb((size_t)nw * o.sc, 0);
    for (int w = 0; w < nw; w++) {
        const uint32_t l = t.words_h[w];
        for (int g = 0; g < o.nlow; g++) {
            const HopDev &h = low[g];
            const uint32_t blo = (l >> h.lo) & 1, bhi = (l >> h.hi) & 1;
            if ((h.nzmask >> (blo | (bhi << 1))) & 1)
                lnb[(size_t)w * o.sl + g] = (uint8_t)t.rank_h[l ^ ((1u << h.lo) | (1u << h.hi))];
        }
        for (int c = 0; c < o.ncross; c++) {
            const HopDev &h = cross[c];
            uint32_t l2 = l ^ (1u << h.lo);
            if (zonly && h.top) l2 = ~l2 & lmask;
            cnb[(size_t)w * o.sc + c] = (uint16_t)(t.rank_h[l2] | (((l >> h.lo) & 1) << 15));
        }
    }
    o.lnb.alloc(lnb.size());
    QS_HIP(hipMemcpy(o.lnb.get(), lnb.data(), lnb.size(), hipMemcpyHostToDevice));
    o.cnb.alloc(cnb.size());
    QS_HIP(hipMemcpy(o.cnb.get(), cnb.data(), cnb.size() * 2, hipMemcpyHostToDevice));
}

void split_classify(const OpHost &op);

}  // namespace

// the lean split kernel applies (split kernel conditions, and the hop lists
// fit its tables)
bool split_lean_ok(const OpHost &op, bool adjoint) {
    if (!split_eligible(op, adjoint)) return false;
    split_classify(op);
    return (adjoint ? op.split_hops_adj : op.split_hops).lean;
}

namespace {

// hop groups of the operator classified by where they sit (once per operator)
void split_classify(const OpHost &op) {
    std::call_once(op.split_once, [&] {
        const SplitBasis &t = split_basis(*op.basis);
        const int H = t.H;
        auto build = [&](const FastTerms &f, SplitHops &out) {
            std::vector<HopDev> low, high, cross;
            for (const HopDev &h : f.hops) (h.hi < H ? low : (h.lo >= H ? high : cross)).push_back(h);
            out.nlow = (int)low.size();
            out.nhigh = (int)high.size();
            out.ncross = (int)cross.size();
            lean_tables(out, low, cross, t, op.basis->mode == MODE_ZONLY);
            low.insert(low.end(), high.begin(), high.end());
            low.insert(low.end(), cross.begin(), cross.end());
            if (!low.empty()) {
                out.d.alloc(low.size());
                QS_HIP(hipMemcpy(out.d.get(), low.data(), low.size() * sizeof(HopDev), hipMemcpyHostToDevice));
            }
        };
        build(op.fast, op.split_hops);
        build(op.fast_adj, op.split_hops_adj);
    });
}

template <typename V, bool CC>
void dispatch_split_cm(const MVArgs &a, const SplitDev &D, int cm, int fewer, size_t lds, hipStream_t s,
                       const HopDev *hops, const HopLaunch &H, const void *x, void *y, int chunk) {
    const V *xv = (const V *)x;
    V *yv = (V *)y;
    auto go = [&](auto kern) {
        // persistent workgroups: as many as stay resident, a multiple of 8
        static int cus = 0;
        if (!cus) {
            int dev = 0;
            QS_HIP(hipGetDevice(&dev));
            QS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        }
        int occ = 0;
        QS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kSpBlock, lds));
        // `fewer` resident workgroups per CU than fit: fewer blocks in flight
        // keep more of the recently read ones in the MALL
        const int64_t want = (int64_t)std::max(1, occ - fewer) * cus;
        const int64_t need = ((int64_t)D.nblocks + kSpBlock / 64 - 1) / (kSpBlock / 64);
        const int grid = (int)std::max<int64_t>(8, std::min(want, need) / 8 * 8);
        SplitDev Dc = D;
        // a tiled order brings its tile length; else four blocks per
        // wavefront per chunk
        Dc.chunk = chunk > 0 ? chunk : std::max(1, (grid / 8) * (kSpBlock / 64) * 4);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kSpBlock), lds, s, a, Dc, hops, H.dcls, H.drest, xv, yv);
    };
    if (cm == CM_ZONLY) go(k_matvec_split<V, CC, CM_ZONLY>);
    else go(k_matvec_split<V, CC, CM_FULL>);
}

// Block orders of the tiled form (see SplitBasis::Order): pass A deals out
// runs of 2^kA consecutive high words, pass B the high words sharing their
// bits below kB (and, under spin inversion, the tile of the inverted bits,
// where the hops that set bit L-1 land).  Tile c goes to XCD c % 8.
const SplitBasis::Order &split_order(const BasisHost &b, const SplitBasis &t, int kA, int kB) {
    std::lock_guard<std::mutex> g(t.order_mutex);
    for (auto &o : t.orders)
        if (o->kA == kA && o->kB == kB) return *o;
    auto o = std::make_unique<SplitBasis::Order>();
    o->kA = kA;
    o->kB = kB;
    const int W = t.W;
    std::vector<uint32_t> blist(t.nblocks), bstart(t.nblocks);
    QS_HIP(hipMemcpy(blist.data(), t.blist.get(), blist.size() * 4, hipMemcpyDeviceToHost));
    QS_HIP(hipMemcpy(bstart.data(), t.bstart.get(), bstart.size() * 4, hipMemcpyDeviceToHost));
    auto emit = [](const std::vector<std::vector<int>> &tiles, const std::vector<uint32_t> &bl,
                   const std::vector<uint32_t> &bs, int &C, int &n, DevBuf<uint32_t> &dl, DevBuf<uint32_t> &ds) {
        C = 1;
        for (auto &v : tiles) C = std::max<int>(C, (int)v.size());
        std::vector<uint32_t> l, st;
        for (auto &v : tiles) {
            if (v.empty()) continue;
            for (int i : v) {
                l.push_back(bl[i]);
                st.push_back(bs[i]);
            }
            for (size_t k = v.size(); k < (size_t)C; k++) {
                l.push_back(0xFFFFFFFFu);
                st.push_back(0);
            }
        }
        n = (int)l.size();
        dl.alloc(l.size());
        ds.alloc(st.size());
        QS_HIP(hipMemcpy(dl.get(), l.data(), l.size() * 4, hipMemcpyHostToDevice));
        QS_HIP(hipMemcpy(ds.get(), st.data(), st.size() * 4, hipMemcpyHostToDevice));
    };
    // pass A: consecutive runs (the block list is ascending in h)
    std::vector<std::vector<int>> ta((size_t)1 << (W - kA));
    for (int i = 0; i < t.nblocks; i++) ta[blist[i] >> kA].push_back(i);
    emit(ta, blist, bstart, o->CA, o->nA, o->blA, o->bsA);
    // pass B: equal low bits m (paired with the inverted bits under spin inversion)
    const uint32_t mm = (1u << kB) - 1;
    std::vector<std::vector<int>> bym((size_t)1 << kB);
    for (int i = 0; i < t.nblocks; i++) bym[blist[i] & mm].push_back(i);
    std::vector<std::vector<int>> tb;
    for (uint32_t m = 0; m <= mm; m++) {
        if (b.mode == MODE_ZONLY) {
            const uint32_t m2 = ~m & mm;
            if (m2 < m) continue;
            std::vector<int> v = bym[m];
            if (m2 != m) v.insert(v.end(), bym[m2].begin(), bym[m2].end());
            tb.push_back(std::move(v));
        } else {
            tb.push_back(bym[m]);
        }
    }
    emit(tb, blist, bstart, o->CB, o->nB, o->blB, o->bsB);
    t.orders.push_back(std::move(o));
    return *t.orders.back();
}

// Hop groups dealt to the two passes for (kA, kB); ok = every group found a pass.
const SplitPassHops &split_pass_hops(const OpHost &op, const SplitBasis &t, int kA, int kB) {
    const int H = t.H;
    std::lock_guard<std::mutex> g(op.pass_mutex);
    for (auto &p : op.pass_hops)
        if (p->kA == kA && p->kB == kB) return *p;
    auto p = std::make_unique<SplitPassHops>();
    p->kA = kA;
    p->kB = kB;
    auto build = [&](const FastTerms &f, SplitHops &A, SplitHops &B) {
        std::vector<HopDev> lowA, highA, crossA, highB, crossB;
        for (const HopDev &h : f.hops) {
            if (h.hi < H + kA) {
                (h.hi < H ? lowA : (h.lo >= H ? highA : crossA)).push_back(h);
            } else if (h.lo < H) {
                crossB.push_back(h);
            } else if (h.lo >= H + kB) {
                highB.push_back(h);
            } else {
                return false;
            }
        }
        auto put = [&](SplitHops &o, std::vector<HopDev> low, const std::vector<HopDev> &high,
                       const std::vector<HopDev> &cross) {
            o.nlow = (int)low.size();
            o.nhigh = (int)high.size();
            o.ncross = (int)cross.size();
            lean_tables(o, low, cross, t, op.basis->mode == MODE_ZONLY);
            low.insert(low.end(), high.begin(), high.end());
            low.insert(low.end(), cross.begin(), cross.end());
            if (!low.empty()) {
                o.d.alloc(low.size());
                QS_HIP(hipMemcpy(o.d.get(), low.data(), low.size() * sizeof(HopDev), hipMemcpyHostToDevice));
            }
        };
        put(A, lowA, highA, crossA);
        put(B, {}, highB, crossB);
        return A.lean && B.lean;
    };
    p->ok = build(op.fast, p->a, p->b);
    p->ok_adj = build(op.fast_adj, p->a_adj, p->b_adj);
    op.pass_hops.push_back(std::move(p));
    return *op.pass_hops.back();
}

// (kA, kB) of the tiled form for tile budget index d: pass A tiles of at
// most 2^(20 + d) bytes of x (1, 2, 4 or 8 MB for d = 0..3, against one
// XCD's 4 MB L2, which also sees y), pass B split as far up as the couplings
// allow.  False when no split fits.
bool split_tiles(const OpHost &op, const SplitBasis &t, size_t esz, int d, int &kA, int &kB) {
    const BasisHost &b = *op.basis;
    const int W = t.W, H = t.H;
    const double rows = (double)b.Ns / std::max(1, t.nblocks);
    const double budget = std::ldexp(1.0, 20 + d);
    kA = 1;
    while (kA + 1 < W && std::ldexp(rows * (double)esz, kA + 1) <= budget) kA++;
    // pass B may start at the lowest high-word bit of a group pass A leaves
    kB = kA;
    for (const FastTerms *f : {&op.fast, &op.fast_adj})
        for (const HopDev &h : f->hops)
            if (h.hi >= H + kA && h.lo >= H) kB = std::min(kB, h.lo - H);
    return kB >= 1 && kA < W;
}

template <typename V, bool CC>
void dispatch_split2_cm(const MVArgs &a, const Split2Dev &D, int cm, size_t lds, hipStream_t s, const HopDev *hops,
                        const HopLaunch &H, const void *x, void *y, int chunk) {
    auto go = [&](auto kern) {
        static int cus = 0;
        if (!cus) {
            int dev = 0;
            QS_HIP(hipGetDevice(&dev));
            QS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        }
        int occ = 0;
        QS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kSpBlock, lds));
        const int64_t want = (int64_t)std::max(1, occ) * cus;
        const int64_t need = ((int64_t)D.nblocks + kSpBlock / 64 - 1) / (kSpBlock / 64);
        const int grid = (int)std::max<int64_t>(8, std::min(want, need) / 8 * 8);
        Split2Dev Dc = D;
        Dc.chunk = chunk > 0 ? chunk : std::max(1, (grid / 8) * (kSpBlock / 64) * 4);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kSpBlock), lds, s, a, Dc, hops, H.dcls, H.drest, (const V *)x,
                           (V *)y);
    };
    if (cm == CM_ZONLY) go(k_matvec_split2<V, CC, CM_ZONLY>);
    else go(k_matvec_split2<V, CC, CM_FULL>);
}

// One product of the split kernel: the first kernel in one pass (fewer =
// 0..2 resident workgroups held back), or the lean kernel in one pass
// (tiled < 0) or in the two-pass tiled form with tile budget index `tiled`.
void launch_split(const OpHost &op, MVArgs &a, int dtype, bool cc, bool adjoint, const void *x, void *y,
                  hipStream_t s, int fewer, int tiled, bool lean) {
    const BasisHost &b = *op.basis;
    const SplitBasis &t = split_basis(b);
    split_classify(op);
    const size_t nw = (size_t)1 << t.H;
    const size_t esz = dtype_size(dtype);
    if (t.maxblock > 64 * kSpRows) fail(QS_EARG, "internal: split block larger than the wavefront's rows");
    const int cm = b.mode == MODE_ZONLY ? CM_ZONLY : CM_FULL;
    HopLaunch HL{nullptr, adjoint ? op.d_dcls_adj.get() : op.d_dcls.get(),
                 adjoint ? op.d_drest_adj.get() : op.d_drest.get()};
    a.ndcls = (int)(adjoint ? op.fast_adj : op.fast).dcls.size();
    a.ndrest = (int)(adjoint ? op.fast_adj : op.fast).drest.size();
    // the first kernel: one launch over the ascending block list
    auto pass1 = [&](const SplitHops &sh) {
        if (sh.nhigh + sh.ncross > kSpMaxHops) fail(QS_EARG, "internal: too many hop groups for the split kernel");
        SplitDev D;
        memset(&D, 0, sizeof D);
        D.H = t.H;
        D.lmask = (1u << t.H) - 1;
        D.wmask = t.W >= 32 ? 0xFFFFFFFFu : ((1u << t.W) - 1);
        D.blist = t.blist.get();
        D.bstart = t.bstart.get();
        D.nblocks = t.nblocks;
        D.boff = t.boff.get();
        D.rank = t.rank.get();
        D.words = t.words.get();
        for (int p = 0; p <= t.H; p++) {
            D.cls_off[p] = t.cls_off[p];
            D.cls_n[p] = t.cls_n[p];
        }
        D.nlow = sh.nlow;
        D.nhigh = sh.nhigh;
        D.ncross = sh.ncross;
        D.own = 1;
        D.xstride = t.maxblock + 1;
        // LDS: shared rank and word tables; per wavefront the hop list, the own
        // block and as many staged cross blocks as fit a 12 KB region
        const size_t list_bytes = ((sizeof(SpHop) * (sh.nhigh + sh.ncross)) + 15) & ~(size_t)15;
        const size_t blk = (size_t)D.xstride * esz;
        const size_t region = 12 * 1024;
        const int fit = (int)(region > list_bytes + blk ? (region - list_bytes - blk) / blk : 0);
        D.nstage = std::min(sh.ncross, fit);
        D.off_rank = 0;
        D.off_words = (int)(nw * 2);
        D.off_wave = (int)((nw * 4 + 15) & ~(size_t)15);
        D.off_xw = (int)list_bytes;
        D.wave_lds = (int)((list_bytes + (size_t)(1 + D.nstage) * blk + 15) & ~(size_t)15);
        const size_t lds = (size_t)D.off_wave + (size_t)(kSpBlock / 64) * D.wave_lds;
        const HopDev *hops = sh.d.get();
        switch (dtype) {
        case QS_F64: dispatch_split_cm<double, false>(a, D, cm, fewer, lds, s, hops, HL, x, y, 0); break;
        case QS_F32: dispatch_split_cm<float, false>(a, D, cm, fewer, lds, s, hops, HL, x, y, 0); break;
        case QS_C128:
            if (cc) dispatch_split_cm<cplx_d, true>(a, D, cm, fewer, lds, s, hops, HL, x, y, 0);
            else dispatch_split_cm<cplx_d, false>(a, D, cm, fewer, lds, s, hops, HL, x, y, 0);
            break;
        case QS_C64:
            if (cc) dispatch_split_cm<cplx_f, true>(a, D, cm, fewer, lds, s, hops, HL, x, y, 0);
            else dispatch_split_cm<cplx_f, false>(a, D, cm, fewer, lds, s, hops, HL, x, y, 0);
            break;
        default: fail(QS_EARG, "unknown dtype");
        }
    };
    // the lean kernel: one launch over block list (bl, bs) of n entries
    auto pass2 = [&](const SplitHops &sh, const uint32_t *bl, const uint32_t *bs, int n, int own, int chunk,
                     const MVArgs &args) {
        Split2Dev D;
        memset(&D, 0, sizeof D);
        D.H = t.H;
        D.lmask = (1u << t.H) - 1;
        D.wmask = t.W >= 32 ? 0xFFFFFFFFu : ((1u << t.W) - 1);
        D.blist = bl;
        D.bstart = bs;
        D.nblocks = n;
        D.own = own;
        D.boff = t.boff.get();
        for (int p = 0; p <= t.H; p++) {
            D.cls_off[p] = t.cls_off[p];
            D.cls_n[p] = t.cls_n[p];
        }
        D.nlow = sh.nlow;
        D.nhigh = sh.nhigh;
        D.ncross = sh.ncross;
        D.sl = sh.sl;
        D.sc = sh.sc;
        D.words = t.words.get();
        D.lnb = sh.lnb.get();
        D.cnb = sh.cnb.get();
        // LDS: word table, low-hop and cross-hop tables, 256 rows per wavefront
        size_t o = 0;
        auto place = [&](size_t bytes) {
            const size_t at = (o + 15) & ~(size_t)15;
            o = at + bytes;
            return (int)at;
        };
        D.off_words = place(nw * 2);
        D.off_lnb = place(own && sh.nlow ? nw * sh.sl : 0);
        D.off_cnb = place(sh.ncross ? nw * sh.sc * 2 : 0);
        D.wave_bytes = (int)(256 * esz);
        D.off_wave = place((size_t)(kSpBlock / 64) * D.wave_bytes);
        const size_t lds = o;
        const HopDev *hops = sh.d.get();
        switch (dtype) {
        case QS_F64: dispatch_split2_cm<double, false>(args, D, cm, lds, s, hops, HL, x, y, chunk); break;
        case QS_F32: dispatch_split2_cm<float, false>(args, D, cm, lds, s, hops, HL, x, y, chunk); break;
        case QS_C128:
            if (cc) dispatch_split2_cm<cplx_d, true>(args, D, cm, lds, s, hops, HL, x, y, chunk);
            else dispatch_split2_cm<cplx_d, false>(args, D, cm, lds, s, hops, HL, x, y, chunk);
            break;
        case QS_C64:
            if (cc) dispatch_split2_cm<cplx_f, true>(args, D, cm, lds, s, hops, HL, x, y, chunk);
            else dispatch_split2_cm<cplx_f, false>(args, D, cm, lds, s, hops, HL, x, y, chunk);
            break;
        default: fail(QS_EARG, "unknown dtype");
        }
    };
    const SplitHops &sh = adjoint ? op.split_hops_adj : op.split_hops;
    if (lean) {
        int kA = 0, kB = 0;
        if (tiled >= 0 && split_tiles(op, t, esz, tiled, kA, kB)) {
            const SplitPassHops &ph = split_pass_hops(op, t, kA, kB);
            if (adjoint ? ph.ok_adj : ph.ok) {
                const SplitBasis::Order &o = split_order(b, t, kA, kB);
                // pass A: y = beta y + alpha (own terms + the groups below H + kA);
                // pass B: y += alpha (the other groups) on the stored y
                pass2(adjoint ? ph.a_adj : ph.a, o.blA.get(), o.bsA.get(), o.nA, 1, o.CA, a);
                MVArgs ab = a;
                ab.bre = 1.0;
                ab.bim = 0.0;
                ab.beta_zero = 0;
                ab.user_diag = nullptr;
                pass2(adjoint ? ph.b_adj : ph.b, o.blB.get(), o.bsB.get(), o.nB, 0, o.CB, ab);
                return;
            }
        }
        if (!sh.lean) fail(QS_EARG, "the lean split kernel does not fit this coupling list");
        pass2(sh, t.blist.get(), t.bstart.get(), t.nblocks, 1, 0, a);
        return;
    }
    pass1(sh);
}

template <bool CC, int CM, bool TRACK, bool PC>
void sell_kernels(const MVArgs &a, const HopLaunch &H, size_t lds, hipStream_t s, int grid, uint32_t *cnt,
                  const SellHost *t) {
    if (!t) {
        hipLaunchKernelGGL((k_sell_count<CC, CM, TRACK, PC>), dim3(grid), dim3(kSellBlock), lds, s, a, H.hops,
                           H.dcls, H.drest, cnt);
    } else {
        hipLaunchKernelGGL((k_sell_fill<CC, CM, TRACK, PC>), dim3(grid), dim3(kSellBlock), lds, s, a, H.hops,
                           H.dcls, H.drest, t->sptr.get(), t->width.get(), t->cols.get(), t->wre.get(),
                           t->cplx ? t->wim.get() : nullptr);
    }
}

void sell_pass(const MVArgs &a, bool cc, int cm, bool track, bool pc, const HopLaunch &H, size_t lds, hipStream_t s,
               int grid, uint32_t *cnt, const SellHost *t) {
#define QS_SELL(CCV, CMV, TR, P) sell_kernels<CCV, CMV, TR, P>(a, H, lds, s, grid, cnt, t)
#define QS_SELL_CM(CCV, TR, P)                                        \
    do {                                                              \
        if (cm == CM_FULL) QS_SELL(CCV, CM_FULL, TR, P);              \
        else if (cm == CM_ZONLY) QS_SELL(CCV, CM_ZONLY, TR, P);       \
        else QS_SELL(CCV, CM_PTAB, TR, P);                            \
    } while (0)
    if (cc) {
        if (pc) { if (track) QS_SELL_CM(true, true, true); else QS_SELL_CM(true, false, true); }
        else { if (track) QS_SELL_CM(true, true, false); else QS_SELL_CM(true, false, false); }
    } else {
        if (pc) { if (track) QS_SELL_CM(false, true, true); else QS_SELL_CM(false, false, true); }
        else { if (track) QS_SELL_CM(false, true, false); else QS_SELL_CM(false, false, false); }
    }
#undef QS_SELL_CM
#undef QS_SELL
}

}  // namespace

// operators.jl:461-576 (_assemble) builds P'.H_par.P as a sparse matrix; here
// the rows of the plain product come straight from the pull kernel's
// neighbour enumeration (two passes: lengths, then entries).
// ----------------------------------------------- weight table of the rows --
// Assembled rows of a symmetry sector hold few distinct weights (couplings
// times characters times orbit-norm ratios), so each stored entry can keep a
// 16-bit index into a table of (re, im) pairs instead of 8 or 16 bytes of
// weight: 6 bytes per entry with its column.  The table is an open-addressed
// hash of the weights' bit patterns (linear probing, bounded); any overflow,
// long probe or 64-bit hash collision leaves the per-entry weights in place.
constexpr uint32_t kDictSlots = 1u << 16;
constexpr int kDictProbe = 512;
constexpr unsigned long long kDictEmpty = ~0ull;

__device__ __forceinline__ unsigned long long dict_key(double re, double im) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(re);
    const unsigned long long b = (unsigned long long)__double_as_longlong(im);
    unsigned long long h = (a ^ (b * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull)) * 0xBF58476D1CE4E5B9ull;
    h ^= h >> 31;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 29;
    return h == kDictEmpty ? h - 1 : h;
}

// slot of every distinct weight; the thread that claims a slot writes its
// pair into the table
__global__ __launch_bounds__(256) void k_dict_insert(const double *__restrict__ re, const double *__restrict__ im,
                                                     uint64_t n, unsigned long long *keys, double *dict,
                                                     uint32_t *status) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (uint64_t)gridDim.x * 256) {
        const double wr = re[e], wi = im ? im[e] : 0.0;
        const unsigned long long h = dict_key(wr, wi);
        uint32_t p = (uint32_t)h & (kDictSlots - 1);
        bool ok = false;
        for (int t = 0; t < kDictProbe; t++) {
            unsigned long long cur = keys[p];
            if (cur == kDictEmpty) {
                cur = atomicCAS(&keys[p], kDictEmpty, h);
                if (cur == kDictEmpty) {
                    dict[2 * p] = wr;
                    dict[2 * p + 1] = wi;
                    atomicAdd(&status[1], 1u);
                    ok = true;
                    break;
                }
            }
            if (cur == h) {
                ok = true;
                break;
            }
            p = (p + 1) & (kDictSlots - 1);
        }
        if (!ok) atomicOr(&status[0], 1u);
    }
}

// each entry's slot, checked against the table's pair bit for bit
__global__ __launch_bounds__(256) void k_dict_index(const double *__restrict__ re, const double *__restrict__ im,
                                                    uint64_t n, const unsigned long long *__restrict__ keys,
                                                    const double *__restrict__ dict, uint16_t *idx,
                                                    uint32_t *status) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (uint64_t)gridDim.x * 256) {
        const double wr = re[e], wi = im ? im[e] : 0.0;
        const unsigned long long h = dict_key(wr, wi);
        uint32_t p = (uint32_t)h & (kDictSlots - 1);
        bool ok = false;
        for (int t = 0; t < kDictProbe; t++) {
            const unsigned long long cur = keys[p];
            if (cur == h) {
                ok = __double_as_longlong(dict[2 * p]) == __double_as_longlong(wr) &&
                     __double_as_longlong(dict[2 * p + 1]) == __double_as_longlong(wi);
                break;
            }
            if (cur == kDictEmpty) break;
            p = (p + 1) & (kDictSlots - 1);
        }
        idx[e] = (uint16_t)p;
        if (!ok) atomicOr(&status[0], 1u);
    }
}

__global__ void k_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// The table form of freshly assembled rows (QS_SELL_TABLE=0 keeps the
// per-entry weights); true when it was kept.
bool sell_compress(SellHost &t, hipStream_t s) {
    const char *env = std::getenv("QS_SELL_TABLE");
    if (env && env[0] == '0') return false;
    if (t.stored == 0) return false;
    SBuf<unsigned long long> keys(kDictSlots, s);
    SBuf<uint32_t> status(2, s);
    DevBuf<double> dict((size_t)2 * kDictSlots);
    const unsigned grid = (unsigned)std::min<uint64_t>(8192, (t.stored + 255) / 256);
    hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(256), 0, s, keys.get(), (uint64_t)kDictSlots, kDictEmpty);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (unsigned long long *)status.get(), (uint64_t)1, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(256), 0, s, (unsigned long long *)dict.get(),
                       (uint64_t)2 * kDictSlots, 0ull);
    const double *im = t.cplx ? t.wim.get() : nullptr;
    hipLaunchKernelGGL(k_dict_insert, dim3(grid), dim3(256), 0, s, t.wre.get(), im, t.stored, keys.get(), dict.get(),
                       status.get());
    QS_HIP(hipGetLastError());
    uint32_t st[2] = {0, 0};
    QS_HIP(hipMemcpyAsync(st, status.get(), sizeof(st), hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    if (st[0]) return false;
    DevBuf<uint16_t> widx((size_t)t.stored);
    hipLaunchKernelGGL(k_dict_index, dim3(grid), dim3(256), 0, s, t.wre.get(), im, t.stored, keys.get(), dict.get(),
                       widx.get(), status.get());
    QS_HIP(hipGetLastError());
    QS_HIP(hipMemcpyAsync(st, status.get(), sizeof(st), hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    if (st[0]) return false;
    t.widx = std::move(widx);
    t.dict = std::move(dict);
    t.ndict = st[1];
    t.wre = DevBuf<double>();
    t.wim = DevBuf<double>();
    return true;
}

void sell_assemble(OpHost &op, hipStream_t s) {
    const BasisHost &b = *op.basis;
    if (!hop_eligible(op, false))
        fail(QS_EARG, "assembly needs two-bit hop couplings, L <= 32 and a parent lookup table");
    MVArgs a = base_args(op, false, 1.0, 0.0, 0.0, 0.0, false, false, nullptr);
    const size_t lds = hop_args(a, op, false, false);
    const bool track = basis_tracks(b);
    const bool cc = op.is_complex != 0;
    const int cm = hop_mode(b);
    const bool pc = b.nup >= 0;
    HopLaunch H{op.d_hops.get(), op.d_dcls.get(), op.d_drest.get()};
    const int64_t n = b.Ns;
    const int grid = (int)((n + kSellBlock - 1) / kSellBlock);
    auto t = std::make_unique<SellHost>();
    t->cplx = cc || b.is_complex;
    t->nslices = (n + 63) / 64;
    {
        // stream-ordered scratch and copies: other streams keep running
        SBuf<uint32_t> cnt((size_t)n, s);
        sell_pass(a, cc, cm, track, pc, H, lds, s, grid, cnt.get(), nullptr);
        QS_HIP(hipGetLastError());
        std::vector<uint32_t> hc((size_t)n);
        QS_HIP(hipMemcpyAsync(hc.data(), cnt.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
        QS_HIP(hipStreamSynchronize(s));
        std::vector<uint64_t> sp((size_t)t->nslices);
        std::vector<uint32_t> w((size_t)t->nslices);
        uint64_t at = 0;
        for (int64_t q = 0; q < t->nslices; q++) {
            uint32_t m = 0;
            for (int64_t r = q * 64; r < std::min<int64_t>(n, q * 64 + 64); r++) {
                m = std::max(m, hc[(size_t)r]);
                t->entries += hc[(size_t)r];
            }
            sp[(size_t)q] = at;
            w[(size_t)q] = m;
            at += (uint64_t)64 * m;
        }
        t->stored = at;
        t->sptr.alloc(sp.size());
        t->width.alloc(w.size());
        QS_HIP(hipMemcpyAsync(t->sptr.get(), sp.data(), sp.size() * 8, hipMemcpyHostToDevice, s));
        QS_HIP(hipMemcpyAsync(t->width.get(), w.data(), w.size() * 4, hipMemcpyHostToDevice, s));
        QS_HIP(hipStreamSynchronize(s));
    }
    t->cols.alloc(std::max<uint64_t>(1, t->stored));
    t->wre.alloc(std::max<uint64_t>(1, t->stored));
    if (t->cplx) t->wim.alloc(std::max<uint64_t>(1, t->stored));
    sell_pass(a, cc, cm, track, pc, H, lds, s, grid, nullptr, t.get());
    QS_HIP(hipGetLastError());
    QS_HIP(hipStreamSynchronize(s));
    sell_compress(*t, s);
    op.sell = std::move(t);
}

void sell_to_csr(const OpHost &op, std::vector<int64_t> &rowptr, std::vector<int64_t> &col,
                 std::vector<double> &val) {
    const SellHost &t = *op.sell;
    const int64_t n = op.basis->Ns;
    std::vector<uint64_t> sp((size_t)t.nslices);
    std::vector<uint32_t> w((size_t)t.nslices), c((size_t)t.stored);
    std::vector<double> re((size_t)t.stored), im(t.cplx ? (size_t)t.stored : 0);
    QS_HIP(hipMemcpy(sp.data(), t.sptr.get(), sp.size() * 8, hipMemcpyDeviceToHost));
    QS_HIP(hipMemcpy(w.data(), t.width.get(), w.size() * 4, hipMemcpyDeviceToHost));
    QS_HIP(hipMemcpy(c.data(), t.cols.get(), c.size() * 4, hipMemcpyDeviceToHost));
    if (t.dict.get()) {
        std::vector<uint16_t> ix((size_t)t.stored);
        std::vector<double> tab((size_t)2 * kDictSlots);
        QS_HIP(hipMemcpy(ix.data(), t.widx.get(), ix.size() * 2, hipMemcpyDeviceToHost));
        QS_HIP(hipMemcpy(tab.data(), t.dict.get(), tab.size() * 8, hipMemcpyDeviceToHost));
        for (size_t e = 0; e < ix.size(); e++) {
            re[e] = tab[2 * (size_t)ix[e]];
            if (t.cplx) im[e] = tab[2 * (size_t)ix[e] + 1];
        }
    } else {
        QS_HIP(hipMemcpy(re.data(), t.wre.get(), re.size() * 8, hipMemcpyDeviceToHost));
        if (t.cplx) QS_HIP(hipMemcpy(im.data(), t.wim.get(), im.size() * 8, hipMemcpyDeviceToHost));
    }
    rowptr.assign((size_t)n + 1, 0);
    col.clear();
    val.clear();
    std::vector<std::pair<uint32_t, cd>> ent;
    for (int64_t r = 0; r < n; r++) {
        const uint64_t q = (uint64_t)r >> 6, base = sp[q] + (r & 63);
        ent.clear();
        for (uint32_t k = 0; k < w[q]; k++) {
            const uint64_t at = base + 64 * (uint64_t)k;
            ent.push_back({c[at], cd(re[at], t.cplx ? im[at] : 0.0)});
        }
        std::stable_sort(ent.begin(), ent.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (size_t i = 0; i < ent.size();) {
            cd sum = 0;
            size_t j = i;
            for (; j < ent.size() && ent[j].first == ent[i].first; j++) sum += ent[j].second;
            // exact zeros are not stored (operators.jl:567-572 drops zero diagonals)
            if (sum != cd(0, 0)) {
                col.push_back(ent[i].first);
                val.push_back(sum.real());
                val.push_back(sum.imag());
            }
            i = j;
        }
        rowptr[(size_t)r + 1] = (int64_t)col.size();
    }
}

void ensure_assembled(const OpHost &op, hipStream_t s) {
    if (!op.want_sell || op.sell) return;
    std::lock_guard<std::mutex> g(op.sell_mutex);
    if (!op.sell) {
        try {
            sell_assemble(const_cast<OpHost &>(op), s);
        } catch (const Error &) {
            op.want_sell = false;   // couplings it cannot assemble: stay matrix-free
        }
    }
}

bool matvec_lanczos(const OpHost &op, int dtype, const void *x, void *y, const void *z, const double *zc, int zmode,
                    double *dot, hipStream_t s) {
    const BasisHost &b = *op.basis;
    if (b.wide || !op.sell || !(op.lanes_per_row == 0 || op.lanes_per_row == QS_VARIANT_SELL)) return false;
    const SellHost &t = *op.sell;
    if (t.cplx && !dtype_complex(dtype)) fail(QS_EARG, "a complex sector needs a complex vector dtype");
    const MVArgs a = base_args(op, false, 1.0, 0.0, 0.0, 0.0, false, false, nullptr);
    const int nb = (int)((b.Ns + kSellBlock - 1) / kSellBlock);
    const SellJob q{nullptr, t.sptr.get(), t.width.get(), t.cols.get(), t.weights(), x, y, 0, nb, z, zc, zmode, dot};
    switch (dtype) {
    case QS_F64: hipLaunchKernelGGL(k_sell_lanczos<double>, dim3(nb), dim3(kSellBlock), 0, s, a, q); break;
    case QS_F32: hipLaunchKernelGGL(k_sell_lanczos<float>, dim3(nb), dim3(kSellBlock), 0, s, a, q); break;
    case QS_C128: hipLaunchKernelGGL(k_sell_lanczos<cplx_d>, dim3(nb), dim3(kSellBlock), 0, s, a, q); break;
    case QS_C64: hipLaunchKernelGGL(k_sell_lanczos<cplx_f>, dim3(nb), dim3(kSellBlock), 0, s, a, q); break;
    default: fail(QS_EARG, "unsupported dtype");
    }
    QS_HIP(hipGetLastError());
    return true;
}

void launch_matvec(const OpHost &op, int dtype, const void *x, void *y, double are, double aim, double bre,
                   double bim, bool adjoint, bool conj_x, bool conj_y, hipStream_t s, double *dot_partial,
                   void *xs_scratch) {
    const BasisHost &b = *op.basis;
    if (b.wide) {
        launch_matvec_wide(op, dtype, x, y, are, aim, bre, bim, adjoint, conj_x, conj_y, s, dot_partial);
        return;
    }
    MVArgs a = base_args(op, adjoint, are, aim, bre, bim, conj_x, conj_y, dot_partial);
    const bool track = basis_tracks(b);
    const bool cc = op.is_complex != 0;
    if (cc && !dtype_complex(dtype)) fail(QS_EARG, "a complex operator needs a complex vector dtype");
    if (xs_scratch) {
        // the no-allocation product: only kernels whose tables exist from
        // qs_op_create (no lazy assembly, no split tables); parent-table
        // sectors pre-scale x into the caller's scratch
        if (hop_eligible(op, adjoint)) {
            const size_t lds = hop_args(a, op, adjoint, false);
            const int cm = hop_mode(b);
            const size_t wsz = dtype_complex(dtype) ? 16 : 8;
            const bool ps = cm == CM_PTAB && (uint64_t)b.Ns * wsz < ((uint64_t)1 << 32) &&
                            (uint64_t)b.ptab.n * 4 < ((uint64_t)1 << 32);
            Launch L{grid_for(b.Ns, kMvBlock), lds, s, nullptr, nullptr, nullptr};
            HopLaunch H{adjoint ? op.d_hops_adj.get() : op.d_hops.get(),
                        adjoint ? op.d_dcls_adj.get() : op.d_dcls.get(),
                        adjoint ? op.d_drest_adj.get() : op.d_drest.get()};
            dispatch_hop(a, dtype, cc, cm, track, 1, ps, x, y, L, H, op, ps ? xs_scratch : nullptr);
            QS_HIP(hipGetLastError());
            return;
        }
    }
    // an operator marked for assembly is assembled by its first plain product,
    // on the stream that product runs on
    if (op.want_sell && !adjoint && !xs_scratch) {
        std::lock_guard<std::mutex> g(op.sell_mutex);
        if (!op.sell) {
            try {
                sell_assemble(const_cast<OpHost &>(op), s);
            } catch (const Error &) {
                op.want_sell = false;   // couplings it cannot assemble: stay matrix-free
            }
        }
    }
    // an assembled operator takes its plain products from the stored rows
    if (op.sell && !adjoint && !dot_partial && !xs_scratch &&
        (op.lanes_per_row == 0 || op.lanes_per_row == QS_VARIANT_SELL)) {
        const SellHost &t = *op.sell;
        if (t.cplx && !dtype_complex(dtype)) fail(QS_EARG, "a complex sector needs a complex vector dtype");
        const int grid = (int)((b.Ns + kSellBlock - 1) / kSellBlock);
        auto go = [&](auto *tag) {
            using V = std::remove_pointer_t<decltype(tag)>;
            if (t.cplx)
                hipLaunchKernelGGL((k_sell_spmv<V, true>), dim3(grid), dim3(kSellBlock), 0, s, a, t.sptr.get(),
                                   t.width.get(), t.cols.get(), t.weights(), (const V *)x, (V *)y);
            else
                hipLaunchKernelGGL((k_sell_spmv<V, false>), dim3(grid), dim3(kSellBlock), 0, s, a, t.sptr.get(),
                                   t.width.get(), t.cols.get(), t.weights(), (const V *)x, (V *)y);
        };
        switch (dtype) {
        case QS_F64: go((double *)nullptr); break;
        case QS_F32: go((float *)nullptr); break;
        case QS_C128: go((cplx_d *)nullptr); break;
        case QS_C64: go((cplx_f *)nullptr); break;
        default: fail(QS_EARG, "unknown dtype");
        }
        QS_HIP(hipGetLastError());
        return;
    }

    // automatic choice: the split kernel (tiled form, 4 MB budget) for large
    // implicit sectors
    const bool want_split = op.lanes_per_row >= QS_VARIANT_SPLIT && op.lanes_per_row <= QS_VARIANT_SPLIT_MAX;
    const bool auto_split = op.lanes_per_row == 0 && b.Ns >= ((int64_t)1 << 20) && split_eligible(op, adjoint);
    if ((want_split || auto_split) && !dot_partial && !xs_scratch) {
        if (!split_eligible(op, adjoint))
            fail(QS_EARG, "the split matvec variant needs two-bit hop couplings on a basis with a particle number "
                          "and no symmetry or spin inversion only, L <= 32");
        const int k = want_split ? op.lanes_per_row - QS_VARIANT_SPLIT
                                 : (split_lean_ok(op, adjoint) ? QS_SPLIT_LEAN : 0);
        if (k < QS_SPLIT_LEAN) launch_split(op, a, dtype, cc, adjoint, x, y, s, k, -1, false);
        else if (k == QS_SPLIT_LEAN) launch_split(op, a, dtype, cc, adjoint, x, y, s, 0, -1, true);
        else launch_split(op, a, dtype, cc, adjoint, x, y, s, 0, k - QS_SPLIT_TILED, true);
        QS_HIP(hipGetLastError());
        return;
    }
    const bool want_hop = op.lanes_per_row >= QS_VARIANT_HOP && op.lanes_per_row <= QS_VARIANT_HOP + 3;
    const bool hop = hop_eligible(op, adjoint) && (op.lanes_per_row == 0 || op.lanes_per_row >= QS_VARIANT_SPLIT ||
                                                   want_hop);   // split / assembled variants fall back here
    if (want_hop && !hop)
        fail(QS_EARG, "the hop matvec variant needs two-bit hop couplings, L <= 32 and a parent lookup table");
    if (hop) {
        const size_t lds = hop_args(a, op, adjoint, dot_partial != nullptr);
        // waves sharing one row tile: QS_VARIANT_HOP + k selects 2^k (k <= 2),
        // + 3 one wave.  Parent-table sectors of 2^17 rows or more (and + 3
        // at any size) gather a pre-scaled x: one gather per neighbour
        // instead of x and 1/sqrt(D), the kernel being bound by the issue of
        // divergent gathers (J1-J2 chain, L=28 k=0 p=+1 z=+1, 361,270 rows:
        // 58 us against 73 us per product, the pre-scale pass included)
        const int cm = hop_mode(b);
        const int k = want_hop ? op.lanes_per_row - QS_VARIANT_HOP : 0;
        const int ws = k == 3 ? 1 : 1 << k;
        const size_t wsz = dtype_complex(dtype) ? 16 : 8;
        const bool ps = cm == CM_PTAB && (k == 3 || b.Ns >= ((int64_t)1 << 17)) &&
                        (uint64_t)b.Ns * wsz < ((uint64_t)1 << 32) && (uint64_t)b.ptab.n * 4 < ((uint64_t)1 << 32);
        const int grid = dot_partial ? matvec_grid(b.Ns) : grid_for(b.Ns, kMvBlock / ws);
        Launch L{grid, lds, s, nullptr, nullptr, nullptr};
        HopLaunch H{adjoint ? op.d_hops_adj.get() : op.d_hops.get(), adjoint ? op.d_dcls_adj.get() : op.d_dcls.get(),
                    adjoint ? op.d_drest_adj.get() : op.d_drest.get()};
        dispatch_hop(a, dtype, cc, cm, track, ws, ps, x, y, L, H, op, nullptr);
        QS_HIP(hipGetLastError());
        return;
    }
    // the general kernels canonicalise with the translation / point group;
    // a general-map basis has only its parent table
    if (b.group.general)
        fail(QS_EARG, "a basis of general symmetry maps takes two-bit hop couplings (the parent-table kernels)");
    int lpr = op.lanes_per_row > 0 ? op.lanes_per_row : default_lanes_per_row(b);
    if (b.mode != MODE_STORED) lpr = 1;
    if (dot_partial) lpr = 8;   // the partial-sum grid is sized by matvec_grid
    // dynamic LDS layout
    size_t off = 0;
    auto place = [&](size_t bytes) {
        size_t o = (off + 15) & ~(size_t)15;
        off = o + bytes;
        return (int)o;
    };
    const bool tlds = b.mode == MODE_STORED && a.ngroups <= kMaxLdsGroups && a.noff <= kMaxLdsTerms;
    a.off_binom = place((b.mode != MODE_STORED && !b.use64 && b.nup >= 0) ? sizeof(uint32_t) * 33 * 65 : 0);
    a.off_grp = place(tlds ? sizeof(GroupOfTerms) * a.ngroups : 0);
    a.off_term = place(tlds ? sizeof(TermDev) * a.noff : 0);
    a.off_red = place(dot_partial ? sizeof(double) * 2 * std::max(kMvBlock, kStBlock) : 0);
    int grid = b.mode == MODE_STORED ? grid_for(b.Ns, kStBlock / lpr) : grid_for(b.Ns, kMvBlock);
    if (dot_partial) grid = matvec_grid(b.Ns);
    const bool adj = adjoint;
    Launch L{grid, off, s, adj ? op.d_groups_adj.get() : op.d_groups.get(), adj ? op.d_off_adj.get() : op.d_off.get(),
             adj ? op.d_diag_adj.get() : op.d_diag.get()};
    if (b.use64) dispatch_v<uint64_t>(a, dtype, cc, b.mode, track, tlds, lpr, x, y, L);
    else dispatch_v<uint32_t>(a, dtype, cc, b.mode, track, tlds, lpr, x, y, L);
    QS_HIP(hipGetLastError());
}

// ------------------------------------------- no-allocation product ----
// The pre-scaled x of qs_op_matvec_noalloc lives in this fixed array, one
// copy per GPU for the life of the process.  In host code its name stands
// for no device address: the address comes from hipGetSymbolAddress, once per
// device with that device current, and only that address is passed on.
static __device__ __align__(256) char scratch[QS_NOALLOC_SCRATCH_BYTES];

void *noalloc_scratch() {
    constexpr int kMaxDevices = 64;
    static std::mutex m;
    static void *addr[kMaxDevices] = {};
    int dev = 0;
    QS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(QS_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> g(m);
    if (!addr[dev]) {
        void *p = nullptr;
        QS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(QS_EDEVICE, "no device address for the no-allocation scratch");
        addr[dev] = p;
    }
    return addr[dev];
}

// ------------------------------------------------------ batched matvec --
struct BatchMatvec {
    int dtype = 0;
    std::vector<const OpHost *> ops;
    DevBuf<MVArgs> args;          // device copies, one per operator
    std::vector<int> nb, cm, pc;
    std::vector<size_t> lds;
    std::vector<char> track, cc;
    JobTable jt;                  // job tables of launches over many assembled operators
};

bool batch_matvec_ok(const std::vector<const OpHost *> &ops, int dtype) {
    for (const OpHost *op : ops)
        if (!hop_eligible(*op, false) || (op->is_complex && !dtype_complex(dtype))) return false;
    return !ops.empty();
}

std::unique_ptr<BatchMatvec, void (*)(BatchMatvec *)> make_batch_matvec(const std::vector<const OpHost *> &ops,
                                                                         int dtype) {
    if (!batch_matvec_ok(ops, dtype))
        fail(QS_EARG, "batched matvec needs two-bit hop couplings (L <= 32, parent lookup tables) for every operator");
    std::unique_ptr<BatchMatvec, void (*)(BatchMatvec *)> bm(new BatchMatvec, [](BatchMatvec *p) { delete p; });
    bm->dtype = dtype;
    bm->ops = ops;
    std::vector<MVArgs> host(ops.size());
    for (size_t i = 0; i < ops.size(); i++) {
        const OpHost &op = *ops[i];
        const BasisHost &b = *op.basis;
        host[i] = base_args(op, false, 1.0, 0.0, 0.0, 0.0, false, false, nullptr);
        bm->lds.push_back(hop_args(host[i], op, false, false));
        bm->nb.push_back(grid_for(b.Ns, kMvBlock));
        bm->cm.push_back(hop_mode(b));
        bm->pc.push_back(b.nup >= 0);
        bm->track.push_back(basis_tracks(b));
        bm->cc.push_back(op.is_complex != 0);
    }
    bm->args.alloc(host.size());
    QS_HIP(hipMemcpy(bm->args.get(), host.data(), sizeof(MVArgs) * host.size(), hipMemcpyHostToDevice));
    return bm;
}

namespace {

template <typename V, bool CC>
void batch_launch_cm(int cm, bool pc, bool track, int total, size_t lds, hipStream_t s, const MVJobs &J) {
#define QS_BHOP(CMV, TR, P) \
    hipLaunchKernelGGL((kb_matvec_hop<V, CC, CMV, TR, P>), dim3(total), dim3(kMvBlock), lds, s, J)
    if (pc) {
        if (cm == CM_FULL) QS_BHOP(CM_FULL, false, true);
        else if (cm == CM_ZONLY) QS_BHOP(CM_ZONLY, false, true);
        else if (track) QS_BHOP(CM_PTAB, true, true);
        else QS_BHOP(CM_PTAB, false, true);
    } else {
        if (cm == CM_FULL) QS_BHOP(CM_FULL, false, false);
        else if (cm == CM_ZONLY) QS_BHOP(CM_ZONLY, false, false);
        else if (track) QS_BHOP(CM_PTAB, true, false);
        else QS_BHOP(CM_PTAB, false, false);
    }
#undef QS_BHOP
}

}  // namespace

void launch_batch_matvec(BatchMatvec &bm, const std::vector<MatvecJob> &jobs, hipStream_t s) {
    // jobs of assembled operators: one launch over their stored rows
    std::vector<char> done(jobs.size(), 0);
    {
        std::vector<size_t> asm_jobs;
        for (size_t j = 0; j < jobs.size(); j++)
            if (bm.ops[jobs[j].op]->sell) asm_jobs.push_back(j);
        if (asm_jobs.size() > (size_t)kSellJobsInline) {
            // one launch through a device job table, each job on nbmax blocks;
            // the jobs of one operator side by side, so its stored rows stay
            // in the L2 / MALL while they run
            std::stable_sort(asm_jobs.begin(), asm_jobs.end(),
                             [&](size_t a, size_t b) { return jobs[a].op < jobs[b].op; });
            std::vector<SellJob> tab;
            int nbmax = 0;
            for (size_t j : asm_jobs) {
                const MatvecJob &mj = jobs[j];
                const OpHost &op = *bm.ops[mj.op];
                const SellHost &t = *op.sell;
                const int nb = (int)((op.basis->Ns + kSellBlock - 1) / kSellBlock);
                tab.push_back(SellJob{bm.args.get() + mj.op, t.sptr.get(), t.width.get(), t.cols.get(), t.weights(),
                                      mj.x, mj.y, 0, nb, mj.z, mj.zc, mj.zmode, mj.dot});
                nbmax = std::max(nbmax, nb);
                done[j] = 1;
            }
            if ((int64_t)nbmax * (int64_t)tab.size() > INT32_MAX) fail(QS_EARG, "internal: batched grid too large");
            const SellJob *d = (const SellJob *)bm.jt.stage(tab.data(), tab.size() * sizeof(SellJob), s);
            const int grid = nbmax * (int)tab.size();
            switch (bm.dtype) {
            case QS_F64: hipLaunchKernelGGL(kb_sell_tab<double>, dim3(grid), dim3(kSellBlock), 0, s, d, nbmax); break;
            case QS_F32: hipLaunchKernelGGL(kb_sell_tab<float>, dim3(grid), dim3(kSellBlock), 0, s, d, nbmax); break;
            case QS_C128: hipLaunchKernelGGL(kb_sell_tab<cplx_d>, dim3(grid), dim3(kSellBlock), 0, s, d, nbmax); break;
            case QS_C64: hipLaunchKernelGGL(kb_sell_tab<cplx_f>, dim3(grid), dim3(kSellBlock), 0, s, d, nbmax); break;
            default: fail(QS_EARG, "unknown dtype");
            }
            QS_HIP(hipGetLastError());
            asm_jobs.clear();
        }
        for (size_t c0 = 0; c0 < asm_jobs.size(); c0 += kSellJobsInline) {
            SellJobs J;
            J.njobs = (int)std::min<size_t>(kSellJobsInline, asm_jobs.size() - c0);
            int total = 0;
            for (int i = 0; i < J.njobs; i++) {
                const MatvecJob &mj = jobs[asm_jobs[c0 + i]];
                const OpHost &op = *bm.ops[mj.op];
                const SellHost &t = *op.sell;
                const int nb = (int)((op.basis->Ns + kSellBlock - 1) / kSellBlock);
                J.job[i] = SellJob{bm.args.get() + mj.op, t.sptr.get(), t.width.get(), t.cols.get(), t.weights(),
                                   mj.x, mj.y, total, nb, mj.z, mj.zc, mj.zmode, mj.dot};
                total += nb;
                done[asm_jobs[c0 + i]] = 1;
            }
            switch (bm.dtype) {
            case QS_F64: hipLaunchKernelGGL(kb_sell_spmv<double>, dim3(total), dim3(kSellBlock), 0, s, J); break;
            case QS_F32: hipLaunchKernelGGL(kb_sell_spmv<float>, dim3(total), dim3(kSellBlock), 0, s, J); break;
            case QS_C128: hipLaunchKernelGGL(kb_sell_spmv<cplx_d>, dim3(total), dim3(kSellBlock), 0, s, J); break;
            case QS_C64: hipLaunchKernelGGL(kb_sell_spmv<cplx_f>, dim3(total), dim3(kSellBlock), 0, s, J); break;
            default: fail(QS_EARG, "unknown dtype");
            }
            QS_HIP(hipGetLastError());
        }
    }
    // the others: one launch per (column mode, particle number) class, jobs in order
    for (size_t first = 0; first < jobs.size(); first++) {
        if (done[first]) continue;
        if (jobs[first].dot || jobs[first].z) fail(QS_EARG, "internal: the fused Lanczos product needs an assembled operator");
        const int key_cm = bm.cm[jobs[first].op], key_pc = bm.pc[jobs[first].op];
        std::vector<size_t> group;
        for (size_t j = first; j < jobs.size(); j++)
            if (!done[j] && bm.cm[jobs[j].op] == key_cm && bm.pc[jobs[j].op] == key_pc) group.push_back(j);
        for (size_t c0 = 0; c0 < group.size(); c0 += kMaxMVJobs) {
            MVJobs J;
            J.njobs = (int)std::min<size_t>(kMaxMVJobs, group.size() - c0);
            int total = 0;
            size_t lds = 0;
            bool track = false, cc = false;
            for (int i = 0; i < J.njobs; i++) {
                const MatvecJob &mj = jobs[group[c0 + i]];
                const OpHost &op = *bm.ops[mj.op];
                J.job[i] = MVJob{bm.args.get() + mj.op, op.d_hops.get(), op.d_dcls.get(), op.d_drest.get(), mj.x, mj.y,
                                 total, bm.nb[mj.op]};
                total += bm.nb[mj.op];
                lds = std::max(lds, bm.lds[mj.op]);
                track = track || bm.track[mj.op];
                cc = cc || bm.cc[mj.op];
                done[group[c0 + i]] = 1;
            }
            switch (bm.dtype) {
            case QS_F64: batch_launch_cm<double, false>(key_cm, key_pc, track, total, lds, s, J); break;
            case QS_F32: batch_launch_cm<float, false>(key_cm, key_pc, track, total, lds, s, J); break;
            case QS_C128:
                if (cc) batch_launch_cm<cplx_d, true>(key_cm, key_pc, track, total, lds, s, J);
                else batch_launch_cm<cplx_d, false>(key_cm, key_pc, track, total, lds, s, J);
                break;
            case QS_C64:
                if (cc) batch_launch_cm<cplx_f, true>(key_cm, key_pc, track, total, lds, s, J);
                else batch_launch_cm<cplx_f, false>(key_cm, key_pc, track, total, lds, s, J);
                break;
            default: fail(QS_EARG, "unknown dtype");
            }
            QS_HIP(hipGetLastError());
        }
    }
}

void launch_matvec_wide(const OpHost &op, int dtype, const void *x, void *y, double are, double aim, double bre,
                        double bim, bool adjoint, bool conj_x, bool conj_y, hipStream_t s, double *dot_partial) {
    const BasisHost &b = *op.basis;
    MVArgs a = base_args(op, adjoint, are, aim, bre, bim, conj_x, conj_y, dot_partial);
    const std::vector<WTerm> &dg = adjoint ? op.wdiag_adj : op.wdiag;
    const std::vector<WTerm> &of = adjoint ? op.woff_adj : op.woff;
    const std::vector<WGroupDev> &gr = adjoint ? op.wgroups_adj : op.wgroups;
    a.ndiag = (int)dg.size();
    a.noff = (int)of.size();
    a.ngroups = (int)gr.size();
    WideArgs w;
    w.reps = b.wreps.get();
    w.g = WGroup{b.group.order(), b.L, b.wperm.get(), b.wc.get(), b.gchi.get()};
    w.grp = adjoint ? op.d_wgroups_adj.get() : op.d_wgroups.get();
    w.off = adjoint ? op.d_woff_adj.get() : op.d_woff.get();
    w.diag = adjoint ? op.d_wdiag_adj.get() : op.d_wdiag.get();
    const bool track = basis_tracks(b);
    const bool cc = op.is_complex != 0;
    if (cc && !dtype_complex(dtype)) fail(QS_EARG, "a complex operator needs a complex vector dtype");
    // one block per 32 rows: the grid of the dot partials (matvec_grid)
    const int grid = matvec_grid(b.Ns);
    auto go = [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        if (cc) {
            if (track) hipLaunchKernelGGL((k_matvec_wide<V, true, true>), dim3(grid), dim3(kStBlock), 0, s, a, w,
                                          (const V *)x, (V *)y);
            else hipLaunchKernelGGL((k_matvec_wide<V, true, false>), dim3(grid), dim3(kStBlock), 0, s, a, w,
                                    (const V *)x, (V *)y);
        } else {
            if (track) hipLaunchKernelGGL((k_matvec_wide<V, false, true>), dim3(grid), dim3(kStBlock), 0, s, a, w,
                                          (const V *)x, (V *)y);
            else hipLaunchKernelGGL((k_matvec_wide<V, false, false>), dim3(grid), dim3(kStBlock), 0, s, a, w,
                                    (const V *)x, (V *)y);
        }
    };
    switch (dtype) {
    case QS_F64: go((double *)nullptr); break;
    case QS_F32: go((float *)nullptr); break;
    case QS_C128: go((cplx_d *)nullptr); break;
    case QS_C64: go((cplx_f *)nullptr); break;
    default: fail(QS_EARG, "unknown dtype");
    }
    QS_HIP(hipGetLastError());
}

}  // namespace qs
