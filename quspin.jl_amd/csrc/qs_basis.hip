// qs_basis.hip -- GPU enumeration of spin-1/2 symmetry sectors.
//
// Reference: SpinBasis1D (basis.jl:622-777) enumerates the nup sector with
// Gosper's hack (basis.jl:229-250), builds an explicit projector stage by
// stage (symmetry_basis.jl:57-213) and orders columns by ascending orbit
// minimum.  Here each lane unranks a chunk of parent states (combinadic),
// walks it with Gosper's successor, keeps a state when it is the minimum of
// its orbit under the whole group and the sector character is trivial on its
// stabiliser, and a two-pass compaction (count, scan, write) emits the kept
// orbit minima in ascending order -- the reference's column order.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "qs_device.h"

namespace qs {

namespace {

constexpr int kChunk = 32;      // parent states per lane
constexpr int kBlock = 256;
constexpr int kSeg = 32;        // search segment (entries read by one scan)

template <typename S>
__device__ __forceinline__ S gosper_next(S s) {
    S lowest = s & (~s + 1);
    S next = s + lowest;
    return next | ((next ^ s) >> (ctz(lowest) + 2));
}

// Is u the orbit minimum with a character that is trivial on its stabiliser?
// Returns the stabiliser size (0 if u is not a kept representative).
template <typename S>
__device__ int keep_rep(const GroupDev &g, S u, S mask) {
    for (int p = 0; p < g.npoint; p++) {
        S v = point_apply(g, p, u);
        int sh = 0;
        for (int t = 0; t < g.nT; t++) {
            if (rotl_bits(v, sh, g.L, mask) < u) return 0;
            sh += g.a;
        }
    }
    int cnt = 0;
    double cre = 0, cim = 0;
    for (int p = 0; p < g.npoint; p++) {
        S v = point_apply(g, p, u);
        int sh = 0;
        for (int t = 0; t < g.nT; t++) {
            if (rotl_bits(v, sh, g.L, mask) == u) {
                cnt++;
                double lr = g.lam_re[t], li = g.lam_im[t];
                double pr = g.pch_re[p], pi = g.pch_im[p];
                cre += lr * pr - li * pi;
                cim += lr * pi + li * pr;
            }
            sh += g.a;
        }
    }
    return (cre * cre + cim * cim > 0.25) ? cnt : 0;
}

template <typename S>
__device__ __forceinline__ S parent_state(int64_t idx, int L, int nup, const uint64_t *binom) {
    if (nup < 0) return (S)idx;
    return comb_unrank<S, uint64_t>(idx, L, nup, binom);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_count(GroupDev g, int nup, int64_t npar, S mask,
                                                  const uint64_t *binom, uint32_t *block_count) {
    __shared__ uint32_t red[kBlock];
    int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    uint32_t cnt = 0;
    if (base < npar) {
        S u = parent_state<S>(base, g.L, nup, binom);
        int64_t end = base + kChunk < npar ? base + kChunk : npar;
        for (int64_t i = base; i < end; i++) {
            if (keep_rep(g, u, mask)) cnt++;
            if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : (S)(u + 1);
        }
    }
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_count[blockIdx.x] = red[0];
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_write(GroupDev g, int nup, int64_t npar, S mask,
                                                  const uint64_t *binom, const uint64_t *block_offset,
                                                  S *reps, uint32_t *orbit, double *inv_nrm) {
    __shared__ uint32_t scan[kBlock];
    int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    uint32_t cnt = 0;
    S u0 = 0;
    int64_t end = base;
    if (base < npar) {
        u0 = parent_state<S>(base, g.L, nup, binom);
        end = base + kChunk < npar ? base + kChunk : npar;
        S u = u0;
        for (int64_t i = base; i < end; i++) {
            if (keep_rep(g, u, mask)) cnt++;
            if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : (S)(u + 1);
        }
    }
    // exclusive scan of lane counts in lane order
    scan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t pos = block_offset[blockIdx.x] + scan[threadIdx.x] - cnt;
    if (cnt == 0) return;
    const int order = g.nT * g.npoint;
    S u = u0;
    for (int64_t i = base; i < end; i++) {
        int st = keep_rep(g, u, mask);
        if (st) {
            uint32_t d = (uint32_t)(order / st);
            reps[pos] = u;
            orbit[pos] = d;
            inv_nrm[pos] = 1.0 / sqrt((double)d);
            pos++;
        }
        if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : (S)(u + 1);
    }
}

// Insert every canonical state into the open-addressing column table.  The
// slot a state lands in depends on insertion order, but a lookup always
// returns the one column holding that state, so results are deterministic.
template <typename S>
__global__ void k_hash_insert(const S *reps, int64_t n, uint64_t *keys, uint32_t *vals, uint64_t hmask,
                              int *max_probe) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const S r = reps[i];
    uint64_t h = state_hash((uint64_t)r) & hmask;
    int probe = 1;
    const unsigned long long empty = ~0ULL;
    if constexpr (sizeof(S) == 4) {
        const unsigned long long ent = ((unsigned long long)r << 32) | (uint32_t)i;
        while (atomicCAS((unsigned long long *)&keys[h], empty, ent) != empty) {
            h = (h + 1) & hmask;
            probe++;
        }
    } else {
        while (atomicCAS((unsigned long long *)&keys[h], empty, (unsigned long long)r) != empty) {
            h = (h + 1) & hmask;
            probe++;
        }
        vals[h] = (uint32_t)i;
    }
    atomicMax(max_probe, probe);
}

// Column states of an implicit basis (every nup-word of the parent range, in
// ascending order): unrank the first state of a chunk, then Gosper.
__global__ __launch_bounds__(kBlock) void k_states(int L, int nup, int64_t n, const uint64_t *binom,
                                                   uint32_t *states) {
    const int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    if (base >= n) return;
    uint32_t u = comb_unrank<uint32_t, uint64_t>(base, L, nup, binom);
    const int64_t end = base + kChunk < n ? base + kChunk : n;
    for (int64_t i = base; i < end; i++) {
        states[i] = u;
        if (i + 1 < end) u = gosper_next(u);
    }
}

// Parent lookup table: entry of parent rank i = column of its orbit minimum
// | (group element reaching the minimum) << colbits, or all ones.  The
// element is the one canonical32 reports, so the table path applies exactly
// the characters of the canonicalising path.
template <bool TRACK>
__global__ __launch_bounds__(kBlock) void k_ptab(GroupDev g, int nup, int64_t npar, const uint64_t *binom,
                                                 const uint64_t *hkeys, uint64_t hmask, int max_probe, int colbits,
                                                 uint32_t *ptab) {
    const int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    if (base >= npar) return;
    uint32_t u = parent_state<uint32_t>(base, g.L, nup, binom);
    const int64_t end = base + kChunk < npar ? base + kChunk : npar;
    for (int64_t i = base; i < end; i++) {
        int bt, bp;
        const uint32_t r = canonical32<TRACK>(g, u, bt, bp);
        const int64_t c = hash_find32(hkeys, hmask, max_probe, r);
        uint32_t e = 0xFFFFFFFFu;
        if (c >= 0) e = (uint32_t)c | (TRACK ? (uint32_t)(bt * g.npoint + bp) << colbits : 0u);
        ptab[i] = e;
        if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : u + 1;
    }
}

// ------------------------------------------------- general site maps ----
// A SpinBasisGeneral group on the device (L <= 32): element e maps u to
// T_e[u & 255] | T_e[256 + (u >> 8 & 255)] | T_e[512 + (u >> 16 & 255)] |
// T_e[768 + (u >> 24)], xor c_e (tab: 1024 words per element), with
// character chi[2 e] + i chi[2 e + 1]; element 0 is the identity.
struct GenGroup {
    int n, L;
    const uint32_t *tab, *c;
    const double *chi;
};

__device__ __forceinline__ uint32_t gen_apply(const GenGroup &g, int e, uint32_t u) {
    const uint32_t *t = g.tab + (size_t)e * 1024;
    return (t[u & 255] | t[256 + ((u >> 8) & 255)] | t[512 + ((u >> 16) & 255)] | t[768 + (u >> 24)]) ^ g.c[e];
}

// the stabiliser size when u is its orbit's minimum and the character is
// trivial on its stabiliser, else 0 (keep_rep for the general group)
__device__ int keep_rep_gen(const GenGroup &g, uint32_t u) {
    int cnt = 0;
    double cre = 0, cim = 0;
    for (int e = 0; e < g.n; e++) {
        const uint32_t w = gen_apply(g, e, u);
        if (w < u) return 0;
        if (w == u) {
            cnt++;
            cre += g.chi[2 * e];
            cim += g.chi[2 * e + 1];
        }
    }
    return (cre * cre + cim * cim > 0.25) ? cnt : 0;
}

__global__ __launch_bounds__(kBlock) void k_count_gen(GenGroup g, int nup, int64_t npar, const uint64_t *binom,
                                                      uint32_t *block_count) {
    __shared__ uint32_t red[kBlock];
    int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    uint32_t cnt = 0;
    if (base < npar) {
        uint32_t u = parent_state<uint32_t>(base, g.L, nup, binom);
        int64_t end = base + kChunk < npar ? base + kChunk : npar;
        for (int64_t i = base; i < end; i++) {
            if (keep_rep_gen(g, u)) cnt++;
            if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : u + 1;
        }
    }
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_count[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_write_gen(GenGroup g, int nup, int64_t npar, const uint64_t *binom,
                                                      const uint64_t *block_offset, uint32_t *reps, uint32_t *orbit,
                                                      double *inv_nrm) {
    __shared__ uint32_t scan[kBlock];
    int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    uint32_t cnt = 0;
    uint32_t u0 = 0;
    int64_t end = base;
    if (base < npar) {
        u0 = parent_state<uint32_t>(base, g.L, nup, binom);
        end = base + kChunk < npar ? base + kChunk : npar;
        uint32_t u = u0;
        for (int64_t i = base; i < end; i++) {
            if (keep_rep_gen(g, u)) cnt++;
            if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : u + 1;
        }
    }
    scan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t pos = block_offset[blockIdx.x] + scan[threadIdx.x] - cnt;
    if (cnt == 0) return;
    uint32_t u = u0;
    for (int64_t i = base; i < end; i++) {
        const int st = keep_rep_gen(g, u);
        if (st) {
            const uint32_t d = (uint32_t)(g.n / st);
            reps[pos] = u;
            orbit[pos] = d;
            inv_nrm[pos] = 1.0 / sqrt((double)d);
            pos++;
        }
        if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : u + 1;
    }
}

// Parent lookup table of a general group: column of the orbit minimum | the
// first element reaching it << colbits (the element canon_host picks too)
__global__ __launch_bounds__(kBlock) void k_ptab_gen(GenGroup g, int nup, int64_t npar, const uint64_t *binom,
                                                     const uint64_t *hkeys, uint64_t hmask, int max_probe,
                                                     int colbits, int track, uint32_t *ptab) {
    const int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kChunk;
    if (base >= npar) return;
    uint32_t u = parent_state<uint32_t>(base, g.L, nup, binom);
    const int64_t end = base + kChunk < npar ? base + kChunk : npar;
    for (int64_t i = base; i < end; i++) {
        uint32_t best = 0xFFFFFFFFu;
        int be = 0;
        for (int e = 0; e < g.n; e++) {
            const uint32_t w = gen_apply(g, e, u);
            if (w < best) {
                best = w;
                be = e;
            }
        }
        const int64_t c = hash_find32(hkeys, hmask, max_probe, best);
        uint32_t ent = 0xFFFFFFFFu;
        if (c >= 0) ent = (uint32_t)c | (track ? (uint32_t)be << colbits : 0u);
        ptab[i] = ent;
        if (i + 1 < end) u = (nup >= 0) ? gosper_next(u) : u + 1;
    }
}

bool group_tracks(const GroupHost &g) {
    if (g.general) {
        for (const cd &c : g.echar)
            if (c != cd(1, 0)) return true;
        return false;
    }
    for (int p = 0; p < g.npoint; p++)
        if (g.pchar[p] != cd(1, 0)) return true;
    for (int t = 0; t < g.nT; t++)
        if (g.lam[t] != cd(1, 0)) return true;
    return false;
}

// Device memory a lookup structure may take: at most a quarter of what is free.
bool fits_device(size_t bytes) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return bytes <= fr / 4;
}

void build_ptab(BasisHost &b) {
    const bool track = group_tracks(b.group);
    const int order = b.group.order();
    int ebits = 0;
    if (track)
        while ((1 << ebits) < order) ebits++;
    const int colbits = 32 - ebits;
    if (b.Ns >= ((int64_t)1 << colbits) - 1) return;
    if (b.npar > ((int64_t)1 << 31) || !fits_device((size_t)b.npar * 4)) return;
    b.ptab.alloc((size_t)b.npar);
    b.ptab_colbits = colbits;
    GroupDev g = group_to_dev(b.group);
    const int64_t lanes = (b.npar + kChunk - 1) / kChunk;
    const unsigned nblk = (unsigned)((lanes + kBlock - 1) / kBlock);
    if (b.group.general) {
        GenGroup gg{order, b.L, b.gtab.get(), b.gc.get(), b.gchi.get()};
        hipLaunchKernelGGL(k_ptab_gen, dim3(nblk), dim3(kBlock), 0, 0, gg, b.nup, b.npar, b.binom.get(),
                           b.hkeys.get(), b.hmask, b.hmax_probe, colbits, (int)track, b.ptab.get());
        QS_HIP(hipGetLastError());
        return;
    }
    if (track)
        hipLaunchKernelGGL((k_ptab<true>), dim3(nblk), dim3(kBlock), 0, 0, g, b.nup, b.npar, b.binom.get(),
                           b.hkeys.get(), b.hmask, b.hmax_probe, colbits, b.ptab.get());
    else
        hipLaunchKernelGGL((k_ptab<false>), dim3(nblk), dim3(kBlock), 0, 0, g, b.nup, b.npar, b.binom.get(),
                           b.hkeys.get(), b.hmask, b.hmax_probe, colbits, b.ptab.get());
    QS_HIP(hipGetLastError());
}

void build_implicit_states(BasisHost &b) {
    if (b.use64 || b.nup < 0 || !fits_device((size_t)b.Ns * 4)) return;
    b.reps32.alloc((size_t)b.Ns);
    const int64_t lanes = (b.Ns + kChunk - 1) / kChunk;
    const unsigned nblk = (unsigned)((lanes + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_states, dim3(nblk), dim3(kBlock), 0, 0, b.L, b.nup, b.Ns, b.binom.get(), b.reps32.get());
    QS_HIP(hipGetLastError());
    b.states_stored = true;
}

// the general group's element tables on the device
void upload_general(BasisHost &b) {
    const GroupHost &h = b.group;
    const int n = h.order();
    std::vector<uint32_t> tab((size_t)n * 1024, 0), c((size_t)n);
    std::vector<double> chi((size_t)2 * n);
    for (int e = 0; e < n; e++) {
        for (int k = 0; k < 4; k++)
            for (int v = 0; v < 256; v++) {
                uint32_t r = 0;
                for (int bit = 0; bit < 8; bit++) {
                    const int i = 8 * k + bit;
                    if (i < b.L && (v >> bit & 1)) r |= 1u << h.eperm[(size_t)e][(size_t)i];
                }
                tab[(size_t)e * 1024 + 256 * k + v] = r;
            }
        c[(size_t)e] = (uint32_t)h.ec[(size_t)e];
        chi[(size_t)2 * e] = h.echar[(size_t)e].real();
        chi[(size_t)2 * e + 1] = h.echar[(size_t)e].imag();
    }
    b.gtab.alloc(tab.size());
    b.gc.alloc(c.size());
    b.gchi.alloc(chi.size());
    QS_HIP(hipMemcpy(b.gtab.get(), tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
    QS_HIP(hipMemcpy(b.gc.get(), c.data(), c.size() * 4, hipMemcpyHostToDevice));
    QS_HIP(hipMemcpy(b.gchi.get(), chi.data(), chi.size() * 8, hipMemcpyHostToDevice));
}

template <typename S>
void build_stored(BasisHost &b, int64_t npar) {
    GroupDev g = group_to_dev(b.group);
    S mask = (S)full_mask(b.L);
    int64_t lanes = (npar + kChunk - 1) / kChunk;
    int64_t nblk = (lanes + kBlock - 1) / kBlock;
    if (nblk > 0x7fffffff) fail(QS_EARG, "sector too large");
    DevBuf<uint32_t> counts((size_t)nblk);
    const bool gen = b.group.general;
    GenGroup gg{b.group.order(), b.L, b.gtab.get(), b.gc.get(), b.gchi.get()};
    if (gen)
        hipLaunchKernelGGL(k_count_gen, dim3((unsigned)nblk), dim3(kBlock), 0, 0, gg, b.nup, npar, b.binom.get(),
                           counts.get());
    else
        hipLaunchKernelGGL((k_count<S>), dim3((unsigned)nblk), dim3(kBlock), 0, 0, g, b.nup, npar, mask,
                           b.binom.get(), counts.get());
    QS_HIP(hipGetLastError());
    std::vector<uint32_t> hc((size_t)nblk);
    QS_HIP(hipMemcpy(hc.data(), counts.get(), sizeof(uint32_t) * nblk, hipMemcpyDeviceToHost));
    std::vector<uint64_t> off((size_t)nblk);
    uint64_t total = 0;
    for (int64_t i = 0; i < nblk; i++) {
        off[i] = total;
        total += hc[i];
    }
    if (total == 0) fail(QS_EARG, "the requested symmetry sector is empty");
    b.Ns = (int64_t)total;
    DevBuf<uint64_t> d_off((size_t)nblk);
    QS_HIP(hipMemcpy(d_off.get(), off.data(), sizeof(uint64_t) * nblk, hipMemcpyHostToDevice));
    // padded to whole 32-entry segments with the largest word, so the
    // search can read any segment with full-width loads
    const size_t padded = (total + kSeg - 1) / kSeg * kSeg;
    DevBuf<S> reps(padded);
    QS_HIP(hipMemset(reps.get(), 0xFF, padded * sizeof(S)));
    b.orbit.alloc((size_t)total);
    b.inv_nrm.alloc((size_t)total);
    if (gen)
        hipLaunchKernelGGL(k_write_gen, dim3((unsigned)nblk), dim3(kBlock), 0, 0, gg, b.nup, npar, b.binom.get(),
                           d_off.get(), (uint32_t *)reps.get(), b.orbit.get(), b.inv_nrm.get());
    else
        hipLaunchKernelGGL((k_write<S>), dim3((unsigned)nblk), dim3(kBlock), 0, 0, g, b.nup, npar, mask,
                           b.binom.get(), d_off.get(), reps.get(), b.orbit.get(), b.inv_nrm.get());
    QS_HIP(hipGetLastError());
    // column hash table: power of two with load factor <= 1/2
    uint64_t hsize = 1024;
    while (hsize < 2 * total) hsize <<= 1;
    b.hmask = hsize - 1;
    b.hkeys.alloc((size_t)hsize);
    QS_HIP(hipMemset(b.hkeys.get(), 0xFF, hsize * sizeof(uint64_t)));
    if (sizeof(S) == 8) b.hvals.alloc((size_t)hsize);
    DevBuf<int> maxp(1);
    QS_HIP(hipMemset(maxp.get(), 0, sizeof(int)));
    hipLaunchKernelGGL((k_hash_insert<S>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, 0, reps.get(),
                       (int64_t)total, b.hkeys.get(), b.hvals.get(), b.hmask, maxp.get());
    QS_HIP(hipGetLastError());
    QS_HIP(hipMemcpy(&b.hmax_probe, maxp.get(), sizeof(int), hipMemcpyDeviceToHost));
    QS_HIP(hipDeviceSynchronize());
    if constexpr (sizeof(S) == 4) b.reps32 = std::move(reps);
    else b.reps64 = std::move(reps);
}

}  // namespace

void build_basis_device(BasisHost &b) {
    // binomial table C(n, k), n < 64, k <= 64, at [n * 65 + k]
    std::vector<uint64_t> tab(64 * 65, 0);
    for (int n = 0; n < 64; n++) {
        tab[n * 65 + 0] = 1;
        for (int k = 1; k <= n; k++) tab[n * 65 + k] = tab[(n - 1) * 65 + k - 1] + (k <= n - 1 ? tab[(n - 1) * 65 + k] : 0);
    }
    b.binom.alloc(tab.size());
    QS_HIP(hipMemcpy(b.binom.get(), tab.data(), sizeof(uint64_t) * tab.size(), hipMemcpyHostToDevice));
    b.use64 = b.L > 32;
    int64_t npar = b.nup >= 0 ? binom(b.L, b.nup) : (int64_t)1 << b.L;
    b.npar = npar;
    if (b.mode == MODE_FULL) {
        b.Ns = npar;
        build_implicit_states(b);
        QS_HIP(hipDeviceSynchronize());
        return;
    }
    if (b.mode == MODE_ZONLY) {
        // canonical states are the ones with bit L-1 clear: C(L-1, nup) of them
        b.Ns = b.nup >= 0 ? binom(b.L - 1, b.nup) : (int64_t)1 << (b.L - 1);
        if (b.Ns == 0) fail(QS_EARG, "the requested symmetry sector is empty");
        build_implicit_states(b);
        QS_HIP(hipDeviceSynchronize());
        return;
    }
    if (b.group.general) upload_general(b);
    if (b.use64) {
        build_stored<uint64_t>(b, npar);
    } else {
        build_stored<uint32_t>(b, npar);
        build_ptab(b);
        QS_HIP(hipDeviceSynchronize());
    }
}

}  // namespace qs
