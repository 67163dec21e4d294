// qs_wide.hip -- bases of general site maps on more than 32 sites (up to 256):
// the reference's SpinBasisGeneral on UInt64 states beyond the 32-bit kernels
// and its WideSpinBasis on UInt256 states (basis.jl:361-445, 530-620).
//
// The parent is the nup sector (ascending, basis.jl:252-280); a parent state
// is kept when it is the minimum of its orbit under the generated group and
// the group's character is trivial on its stabiliser; the kept states in
// ascending order are the columns.  Each thread walks kWChunk consecutive
// parent ranks: the first unranked from the colexicographic rank (saturating
// binomial table), the rest by the next-combination step.
#include "qs_device.h"

#include <algorithm>
#include <vector>

namespace qs {
namespace {

constexpr int kWBlock = 256;
constexpr int kWChunk = 16;
constexpr int kWBin = kMaxWideL + 1;   // binomial table row length

// state of colexicographic rank r among nup-particle states:
// r = sum_k C(c_k, k) over the occupied sites c_nup > ... > c_1
__device__ W4 w4_unrank(int64_t r, int L, int nup, const uint64_t *binom) {
    W4 s{{0, 0, 0, 0}};
    uint64_t rem = (uint64_t)r;
    int c = L - 1;
    for (int k = nup; k >= 1; k--) {
        while (binom[(size_t)c * kWBin + k] > rem) c--;
        s.w[c >> 6] |= 1ULL << (c & 63);
        rem -= binom[(size_t)c * kWBin + k];
        c--;
    }
    return s;
}

// the next larger state with the same number of particles
__device__ W4 w4_next(const W4 &u) {
    W4 r = u;
    int p = 0;
    while (!((r.w[p >> 6] >> (p & 63)) & 1)) p++;
    int q = p;
    while ((r.w[q >> 6] >> (q & 63)) & 1) q++;
    for (int i = p; i < q; i++) r.w[i >> 6] &= ~(1ULL << (i & 63));
    r.w[q >> 6] |= 1ULL << (q & 63);
    for (int i = 0; i < q - p - 1; i++) r.w[i >> 6] |= 1ULL << (i & 63);
    return r;
}

// the stabiliser size when u is its orbit's minimum and the character is
// trivial on its stabiliser, else 0
__device__ int w4_keep(const WGroup &g, const W4 &u) {
    int cnt = 0;
    double cre = 0, cim = 0;
    for (int e = 0; e < g.n; e++) {
        const W4 w = wg_apply(g, e, u);
        if (w4_less(w, u)) return 0;
        if (w4_eq(w, u)) {
            cnt++;
            cre += g.chi[2 * e];
            cim += g.chi[2 * e + 1];
        }
    }
    return (cre * cre + cim * cim > 0.25) ? cnt : 0;
}

__global__ __launch_bounds__(kWBlock) void k_wcount(WGroup g, int nup, int64_t npar, const uint64_t *binom,
                                                    uint32_t *block_count) {
    __shared__ uint32_t red[kWBlock];
    const int64_t base = ((int64_t)blockIdx.x * kWBlock + threadIdx.x) * kWChunk;
    uint32_t cnt = 0;
    if (base < npar) {
        W4 u = w4_unrank(base, g.L, nup, binom);
        const int64_t end = base + kWChunk < npar ? base + kWChunk : npar;
        for (int64_t i = base; i < end; i++) {
            if (w4_keep(g, u)) cnt++;
            if (i + 1 < end) u = w4_next(u);
        }
    }
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = kWBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_count[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kWBlock) void k_wwrite(WGroup g, int nup, int64_t npar, const uint64_t *binom,
                                                    const uint64_t *block_offset, uint64_t *reps, uint32_t *orbit,
                                                    double *inv_nrm) {
    __shared__ uint32_t scan[kWBlock];
    const int64_t base = ((int64_t)blockIdx.x * kWBlock + threadIdx.x) * kWChunk;
    uint32_t cnt = 0;
    W4 u0{{0, 0, 0, 0}};
    int64_t end = base;
    if (base < npar) {
        u0 = w4_unrank(base, g.L, nup, binom);
        end = base + kWChunk < npar ? base + kWChunk : npar;
        W4 u = u0;
        for (int64_t i = base; i < end; i++) {
            if (w4_keep(g, u)) cnt++;
            if (i + 1 < end) u = w4_next(u);
        }
    }
    scan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kWBlock; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t pos = block_offset[blockIdx.x] + scan[threadIdx.x] - cnt;
    if (cnt == 0) return;
    W4 u = u0;
    for (int64_t i = base; i < end; i++) {
        const int st = w4_keep(g, u);
        if (st) {
            const uint32_t d = (uint32_t)(g.n / st);
            for (int k = 0; k < 4; k++) reps[4 * pos + k] = u.w[k];
            orbit[pos] = d;
            inv_nrm[pos] = 1.0 / sqrt((double)d);
            pos++;
        }
        if (i + 1 < end) u = w4_next(u);
    }
}

}  // namespace

void wide_enumerate(BasisHost &b) {
    const GroupHost &h = b.group;
    const int n = h.order(), L = b.L;
    // the group on the device: site tables, xor constants, characters
    std::vector<uint8_t> perm((size_t)n * L);
    std::vector<uint64_t> c((size_t)4 * n);
    std::vector<double> chi((size_t)2 * n);
    for (int e = 0; e < n; e++) {
        for (int i = 0; i < L; i++) perm[(size_t)e * L + i] = (uint8_t)h.eperm[(size_t)e][(size_t)i];
        for (int k = 0; k < 4; k++) c[(size_t)4 * e + k] = h.ecw[(size_t)e].w[k];
        chi[(size_t)2 * e] = h.echar[(size_t)e].real();
        chi[(size_t)2 * e + 1] = h.echar[(size_t)e].imag();
    }
    b.wperm.alloc(perm.size());
    b.wc.alloc(c.size());
    b.gchi.alloc(chi.size());
    QS_HIP(hipMemcpy(b.wperm.get(), perm.data(), perm.size(), hipMemcpyHostToDevice));
    QS_HIP(hipMemcpy(b.wc.get(), c.data(), c.size() * 8, hipMemcpyHostToDevice));
    QS_HIP(hipMemcpy(b.gchi.get(), chi.data(), chi.size() * 8, hipMemcpyHostToDevice));
    // binomial table C(m, k), m, k <= 256, saturating at 2^64 - 1 (unranking
    // compares entries with ranks below C(L, nup), which fits)
    std::vector<uint64_t> tab((size_t)kWBin * kWBin, 0);
    for (int m = 0; m < kWBin; m++) {
        tab[(size_t)m * kWBin] = 1;
        for (int k = 1; k <= m; k++) {
            const uint64_t a = tab[(size_t)(m - 1) * kWBin + k - 1], bb = tab[(size_t)(m - 1) * kWBin + k];
            tab[(size_t)m * kWBin + k] = a > ~0ULL - bb ? ~0ULL : a + bb;
        }
    }
    const uint64_t np = tab[(size_t)L * kWBin + b.nup];
    if (np > ((uint64_t)1 << 40)) fail(QS_EARG, "the selected particle sector is too large");
    b.npar = (int64_t)np;
    DevBuf<uint64_t> d_binom(tab.size());
    QS_HIP(hipMemcpy(d_binom.get(), tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
    WGroup g{n, L, b.wperm.get(), b.wc.get(), b.gchi.get()};
    const int64_t lanes = (b.npar + kWChunk - 1) / kWChunk;
    const int64_t nblk = (lanes + kWBlock - 1) / kWBlock;
    if (nblk > 0x7fffffff) fail(QS_EARG, "sector too large");
    DevBuf<uint32_t> counts((size_t)nblk);
    hipLaunchKernelGGL(k_wcount, dim3((unsigned)nblk), dim3(kWBlock), 0, 0, g, b.nup, b.npar, d_binom.get(),
                       counts.get());
    QS_HIP(hipGetLastError());
    std::vector<uint32_t> hc((size_t)nblk);
    QS_HIP(hipMemcpy(hc.data(), counts.get(), sizeof(uint32_t) * nblk, hipMemcpyDeviceToHost));
    std::vector<uint64_t> off((size_t)nblk);
    uint64_t total = 0;
    for (int64_t i = 0; i < nblk; i++) {
        off[(size_t)i] = total;
        total += hc[(size_t)i];
    }
    if (total == 0) fail(QS_EARG, "the requested symmetry sector is empty");
    b.Ns = (int64_t)total;
    DevBuf<uint64_t> d_off((size_t)nblk);
    QS_HIP(hipMemcpy(d_off.get(), off.data(), sizeof(uint64_t) * nblk, hipMemcpyHostToDevice));
    b.wreps.alloc((size_t)4 * total);
    b.orbit.alloc((size_t)total);
    b.inv_nrm.alloc((size_t)total);
    hipLaunchKernelGGL(k_wwrite, dim3((unsigned)nblk), dim3(kWBlock), 0, 0, g, b.nup, b.npar, d_binom.get(),
                       d_off.get(), b.wreps.get(), b.orbit.get(), b.inv_nrm.get());
    QS_HIP(hipGetLastError());
    QS_HIP(hipDeviceSynchronize());
}

void build_wide_basis(BasisHost &b) {
    if (b.L < 1 || b.L > kMaxWideL) fail(QS_EARG, "L must be between 1 and 256");
    if (b.nup < 0) fail(QS_EARG, "wide spin bases require nup, Nup, or m to avoid enumerating 2^L states");
    build_group_wide(b);
    b.wide = 1;
    b.use64 = 1;
    b.mode = MODE_STORED;
    wide_enumerate(b);
}

}  // namespace qs
