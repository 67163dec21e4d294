// qs_device.h -- device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "qs_internal.h"

namespace qs {

// ---------------------------------------------------------- value types ---
struct cplx_d { double re, im; };
struct cplx_f { float re, im; };

template <typename V> struct VTraits;
template <> struct VTraits<double> {
    static constexpr bool cplx = false;
    __device__ static inline void load(const double *p, int64_t i, double &re, double &im) { re = p[i]; im = 0; }
    __device__ static inline void store(double *p, int64_t i, double re, double) { p[i] = re; }
};
template <> struct VTraits<float> {
    static constexpr bool cplx = false;
    __device__ static inline void load(const float *p, int64_t i, double &re, double &im) { re = (double)p[i]; im = 0; }
    __device__ static inline void store(float *p, int64_t i, double re, double) { p[i] = (float)re; }
};
template <> struct VTraits<cplx_d> {
    static constexpr bool cplx = true;
    __device__ static inline void load(const cplx_d *p, int64_t i, double &re, double &im) {
        cplx_d v = p[i];
        re = v.re;
        im = v.im;
    }
    __device__ static inline void store(cplx_d *p, int64_t i, double re, double im) { p[i] = cplx_d{re, im}; }
};
template <> struct VTraits<cplx_f> {
    static constexpr bool cplx = true;
    __device__ static inline void load(const cplx_f *p, int64_t i, double &re, double &im) {
        cplx_f v = p[i];
        re = (double)v.re;
        im = (double)v.im;
    }
    __device__ static inline void store(cplx_f *p, int64_t i, double re, double im) {
        p[i] = cplx_f{(float)re, (float)im};
    }
};

// ------------------------------------------------------------ bit helpers --
__device__ __forceinline__ uint32_t rev_bits(uint32_t u, int L) { return __builtin_bitreverse32(u) >> (32 - L); }
__device__ __forceinline__ uint64_t rev_bits(uint64_t u, int L) { return __builtin_bitreverse64(u) >> (64 - L); }
__device__ __forceinline__ int popc(uint32_t u) { return __builtin_popcount(u); }
__device__ __forceinline__ int popc(uint64_t u) { return __builtin_popcountll(u); }
__device__ __forceinline__ int ctz(uint32_t u) { return __builtin_ctz(u); }
__device__ __forceinline__ int ctz(uint64_t u) { return __builtin_ctzll(u); }

// rotl within L bits; k in [0, L)
__device__ __forceinline__ uint32_t rotl_bits(uint32_t u, int k, int L, uint32_t mask) {
    uint64_t dd = ((uint64_t)u << L) | u;
    return (uint32_t)(dd >> (L - k)) & mask;
}
__device__ __forceinline__ uint64_t rotl_bits(uint64_t u, int k, int L, uint64_t mask) {
    return ((u << k) | (u >> (L - k))) & mask;
}

template <typename S>
__device__ __forceinline__ S point_apply(const GroupDev &g, int p, S u) {
    S v = g.prev[p] ? rev_bits(u, g.L) : u;
    return v ^ (S)g.pflip[p];
}

// Orbit minimum of u under {T^t o X_p}; returns the (t, p) reaching it.
template <typename S, bool TRACK>
__device__ __forceinline__ S canonical(const GroupDev &g, S u, S mask, int &bt, int &bp) {
    S best = (S)~(S)0;
    bt = 0;
    bp = 0;
    for (int p = 0; p < g.npoint; p++) {
        S v = point_apply(g, p, u);
        int sh = 0;
        for (int t = 0; t < g.nT; t++) {
            S w = rotl_bits(v, sh, g.L, mask);
            if (TRACK) {
                if (w < best) { best = w; bt = t; bp = p; }
            } else {
                best = w < best ? w : best;
            }
            sh += g.a;
        }
    }
    return best;
}

// L <= 32: the same minimum from 32-bit funnel shifts.  P is the state
// repeated periodically over 64 bits; the window (P >> s)[31:0] carries a
// rotation R of the state in its top L bits and, below it, bits that are a
// function of R alone (periodicity).  Comparing windows therefore orders the
// rotations, at one v_alignbit_b32 and one min per group element.
template <bool TRACK>
__device__ __forceinline__ uint32_t canonical32(const GroupDev &g, uint32_t u, int &bt, int &bp) {
    uint32_t best = 0xFFFFFFFFu;
    int bs = 0;
    bp = 0;
    const int L = g.L, a = g.a, nT = g.nT;
    // window offset of T^0; each further translation moves it down by a
    // (mod L).  Kept in scalar registers: no memory access inside the loops.
    const int s0 = g.win[0];
    for (int p = 0; p < g.npoint; p++) {
        uint32_t v = point_apply(g, p, u);
        uint64_t P = v;
        for (int k = L; k < 64; k += L) P |= (uint64_t)v << k;
        const uint32_t lo = (uint32_t)P, hi = (uint32_t)(P >> 32);
        if (a == 1 && nT == L) {
            // every offset 0..L-1 is some translation: scan them in order
#pragma unroll 8
            for (int s = 0; s < L; s++) {
                const uint32_t w = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)s);
                if (TRACK) {
                    if (w < best) { best = w; bs = s; bp = p; }
                } else {
                    best = w < best ? w : best;
                }
            }
        } else {
            int s = s0;
            for (int t = 0; t < nT; t++) {
                const uint32_t w = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)s);
                if (TRACK) {
                    if (w < best) { best = w; bs = s; bp = p; }
                } else {
                    best = w < best ? w : best;
                }
                s -= a;
                if (s < 0) s += L;
            }
        }
    }
    // offset s belongs to T^t with s = s0 - t*a (mod L) and t*a < L
    int d = s0 - bs;
    if (d < 0) d += L;
    bt = d / a;
    return best >> (32 - L);
}

// canonical32 for NB states at once, point operations outermost, so each
// point operation's parameters are read once per batch and the NB
// independent min-chains interleave.  act[j] false leaves best[j] untouched.
template <int NB, bool TRACK>
__device__ __forceinline__ void canonical32_batch(const GroupDev &g, const uint32_t (&u)[NB], uint32_t (&best)[NB],
                                                  int (&bt)[NB], int (&bp)[NB]) {
    const int L = g.L, a = g.a, nT = g.nT;
    const int s0 = g.win[0];
    int bs[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        best[j] = 0xFFFFFFFFu;
        bs[j] = 0;
        bp[j] = 0;
    }
    for (int p = 0; p < g.npoint; p++) {
        const int rev = g.prev[p];
        const uint32_t flip = (uint32_t)g.pflip[p];
        uint32_t lo[NB], hi[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const uint32_t v = (rev ? rev_bits(u[j], L) : u[j]) ^ flip;
            uint64_t P = v;
            for (int k = L; k < 64; k += L) P |= (uint64_t)v << k;
            lo[j] = (uint32_t)P;
            hi[j] = (uint32_t)(P >> 32);
        }
        if (a == 1 && nT == L) {
#pragma unroll 4
            for (int s = 0; s < L; s++) {
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const uint32_t w = __builtin_amdgcn_alignbit(hi[j], lo[j], (uint32_t)s);
                    if (TRACK) {
                        if (w < best[j]) { best[j] = w; bs[j] = s; bp[j] = p; }
                    } else {
                        best[j] = w < best[j] ? w : best[j];
                    }
                }
            }
        } else {
            int s = s0;
            for (int t = 0; t < nT; t++) {
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const uint32_t w = __builtin_amdgcn_alignbit(hi[j], lo[j], (uint32_t)s);
                    if (TRACK) {
                        if (w < best[j]) { best[j] = w; bs[j] = s; bp[j] = p; }
                    } else {
                        best[j] = w < best[j] ? w : best[j];
                    }
                }
                s -= a;
                if (s < 0) s += L;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
        int d = s0 - bs[j];
        if (d < 0) d += L;
        bt[j] = d / a;
        best[j] >>= (32 - L);
    }
}

// Hash of a canonical state for the column table (Fibonacci hashing).
__device__ __forceinline__ uint64_t state_hash(uint64_t r) { return (r * 0x9E3779B97F4A7C15ULL) >> 20; }

// Column of canonical state r (L <= 32) in the (state << 32 | column) table,
// or -1.
__device__ __forceinline__ int64_t hash_find32(const uint64_t *keys, uint64_t hmask, int max_probe, uint32_t r) {
    uint64_t h = state_hash((uint64_t)r) & hmask;
    for (int p = 0; p < max_probe; p++) {
        const uint64_t e = keys[h];
        if (e == ~0ULL) return -1;
        if ((uint32_t)(e >> 32) == r) return (int64_t)(uint32_t)e;
        h = (h + 1) & hmask;
    }
    return -1;
}

// Colex rank from a 32-bit binomial table C(n, k) at [n * 65 + k].
__device__ __forceinline__ uint32_t comb_rank32(uint32_t u, const uint32_t *binom) {
    uint32_t r = 0;
    int k = 1;
    while (u) {
        r += binom[ctz(u) * 65 + k];
        k++;
        u &= u - 1;
    }
    return r;
}

// Colex rank change (mod 2^32) when the particle on bit i or j (i < j, the
// two bits differ) hops to the other one: the moved particle changes index
// and the particles strictly between the two bits shift by one.
__device__ __forceinline__ uint32_t rank_delta32(uint32_t s, int i, int j, const uint32_t *binom) {
    const int c = popc(s & ((1u << i) - 1));
    const uint32_t between = s & ((j < 32 ? (1u << j) : 0u) - (2u << i));
    const int m = popc(between);
    uint32_t d = binom[j * 65 + c + m + 1] - binom[i * 65 + c + 1];
    uint32_t b = between;
    int l = 1;
    while (b) {
        const int q = ctz(b);
        d += binom[q * 65 + c + l] - binom[q * 65 + c + l + 1];
        l++;
        b &= b - 1;
    }
    return ((s >> i) & 1) ? d : (uint32_t)0 - d;
}

// Combinadic (colex) rank of a word with w set bits: sum_k C(p_k, k).
// binom is laid out as C(n, k) at [n * 65 + k].
template <typename S, typename B>
__device__ __forceinline__ int64_t comb_rank(S u, const B *binom) {
    int64_t r = 0;
    int k = 1;
    while (u) {
        int p = ctz(u);
        r += (int64_t)binom[p * 65 + k];
        k++;
        u &= u - 1;
    }
    return r;
}

template <typename S, typename B>
__device__ __forceinline__ S comb_unrank(int64_t i, int L, int w, const B *binom) {
    S u = 0;
    int p = L - 1;
    for (int k = w; k >= 1; k--) {
        while ((int64_t)binom[p * 65 + k] > i) p--;
        u |= (S)1 << p;
        i -= (int64_t)binom[p * 65 + k];
        p--;
    }
    return u;
}

// ---------------------------------------------------------- wide states --
// States of up to 256 sites as four words (site i = bit i % 64 of word i / 64),
// ordered as 256-bit integers.
struct W4 { uint64_t w[4]; };

__device__ __forceinline__ bool w4_less(const W4 &a, const W4 &b) {
    if (a.w[3] != b.w[3]) return a.w[3] < b.w[3];
    if (a.w[2] != b.w[2]) return a.w[2] < b.w[2];
    if (a.w[1] != b.w[1]) return a.w[1] < b.w[1];
    return a.w[0] < b.w[0];
}
__device__ __forceinline__ bool w4_eq(const W4 &a, const W4 &b) {
    return a.w[0] == b.w[0] && a.w[1] == b.w[1] && a.w[2] == b.w[2] && a.w[3] == b.w[3];
}
__device__ __forceinline__ W4 w4_at(const uint64_t *p, int64_t i) {
    const uint64_t *q = p + 4 * i;
    return W4{{q[0], q[1], q[2], q[3]}};
}

// A wide group: element e sends site i to site perm[e L + i], then xors the
// four words c[4 e ..]; character chi[2 e] + i chi[2 e + 1]; element 0 is
// the identity.
struct WGroup {
    int n, L;
    const uint8_t *perm;
    const uint64_t *c;
    const double *chi;
};

__device__ __forceinline__ W4 wg_apply(const WGroup &g, int e, const W4 &u) {
    const uint8_t *pe = g.perm + (size_t)e * g.L;
    W4 r = w4_at(g.c, e);
    for (int k = 0; k < 4; k++) {
        uint64_t bits = u.w[k];
        while (bits) {
            const int t = pe[64 * k + __builtin_ctzll(bits)];
            r.w[t >> 6] ^= 1ULL << (t & 63);
            bits &= bits - 1;
        }
    }
    return r;
}

// Index of s in the ascending array st[0 .. n), or -1.
__device__ __forceinline__ int64_t w4_find(const uint64_t *st, int64_t n, const W4 &s) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (w4_less(w4_at(st, mid), s)) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && w4_eq(w4_at(st, lo), s)) ? lo : -1;
}

// ------------------------------------------------------- assembled rows --
// One assembled row: sum over its w stored entries (column c, weight e) of
// e x[c], entries at base + 64 k.  Groups of kSellU entries load their
// columns and weights first and then gather x, so a lane keeps kSellU
// gathers in flight; the sum runs in entry order either way.
constexpr int kSellU = 4;
template <typename V, bool D>
__device__ __forceinline__ void sell_row_w(const uint32_t *__restrict__ cols, const SellW &W, uint64_t base,
                                           uint32_t w, const V *__restrict__ x, int conj_x, double &ar, double &ai) {
    using T = VTraits<V>;
    auto weight = [&](uint64_t at, double &er, double &ei) {
        if (D) {
            const uint32_t k = W.idx[at];
            er = W.dict[2 * k];
            ei = W.dict[2 * k + 1];
        } else {
            er = W.re[at];
            ei = W.im ? W.im[at] : 0.0;
        }
    };
    uint32_t k = 0;
    for (; k + kSellU <= w; k += kSellU) {
        uint32_t c[kSellU];
        double er[kSellU], ei[kSellU], vr[kSellU], vi[kSellU];
#pragma unroll
        for (int u = 0; u < kSellU; u++) {
            const uint64_t at = base + 64 * (uint64_t)(k + u);
            c[u] = cols[at];
            weight(at, er[u], ei[u]);
        }
#pragma unroll
        for (int u = 0; u < kSellU; u++) T::load(x, c[u], vr[u], vi[u]);
#pragma unroll
        for (int u = 0; u < kSellU; u++) {
            if (conj_x) vi[u] = -vi[u];
            ar += er[u] * vr[u] - ei[u] * vi[u];
            ai += er[u] * vi[u] + ei[u] * vr[u];
        }
    }
    for (; k < w; k++) {
        const uint64_t at = base + 64 * (uint64_t)k;
        const uint32_t c = cols[at];
        double er, ei, vr, vi;
        weight(at, er, ei);
        T::load(x, c, vr, vi);
        if (conj_x) vi = -vi;
        ar += er * vr - ei * vi;
        ai += er * vi + ei * vr;
    }
}

template <typename V>
__device__ __forceinline__ void sell_row(const uint32_t *__restrict__ cols, const SellW &W, uint64_t base, uint32_t w,
                                         const V *__restrict__ x, int conj_x, double &ar, double &ai) {
    if (W.dict) sell_row_w<V, true>(cols, W, base, w, x, conj_x, ar, ai);
    else sell_row_w<V, false>(cols, W, base, w, x, conj_x, ar, ai);
}

// Coefficient of an axpby read from device memory (a reduction result still
// on the device), computed with the same IEEE operations the host would use.
__device__ __forceinline__ void dev_coef(const double *c, int mode, double &ar, double &ai) {
    switch (mode) {
    case CO_NEG: ar = -c[0]; ai = -c[1]; break;
    case CO_NEG_RE: ar = -c[0]; ai = 0.0; break;
    case CO_NEG_SQRT: ar = -sqrt(c[0]); ai = 0.0; break;
    case CO_INV_SQRT: ar = 1.0 / sqrt(c[0]); ai = 0.0; break;
    default: break;
    }
}

}  // namespace qs
