// qs_project.hip -- moving vectors between a sector and its parent space.
//
// Reference: the block projectors of BlockOps (tools.jl:2447-2456 projects
// psi_0 with projector' * psi_0; :2488-2495 accumulates projector * states)
// built by projection_matrix (symmetry_basis.jl:234-255).  Column c of the
// projector is |c> = sum over its orbit of P[u, c] |u> with
//   P[u, c] = chi(h_u) / sqrt(D_c),   h_u u = s_c (the orbit minimum),
// the amplitude qs_basis_lookup returns.  The parent vector is indexed by
// the basis' own parent (kind 0: its particle sector in ascending order,
// index = colex rank; all 2^L states without nup) or by the state itself
// (kind 1: a parent basis without nup around particle-number blocks).
//  * project: out[c] = sum_u conj(P[u, c]) psi[u], gathered per column over
//    the whole group: sum_g chi(g) psi[g s_c] * sqrt(D_c) / |G| (every orbit
//    state appears |G| / D_c times).  No atomics.
//  * expand: parent[u] (+)= P[u, c(u)] x[c(u)] per parent state, from the
//    parent lookup table (or the closed forms of the implicit sectors).
#include <algorithm>

#include "qs_device.h"

namespace qs {

namespace {

constexpr int kPB = 256;

__device__ __forceinline__ uint32_t parent_index(uint32_t u, int nup, int kind, const uint32_t *binom) {
    return (nup >= 0 && kind == 0) ? comb_rank32(u, binom) : u;
}

// parent-vector index of the basis' own parent entry r
__device__ __forceinline__ int64_t own_to_parent(int64_t r, int L, int nup, int kind, const uint64_t *binom) {
    if (kind == 0 || nup < 0) return r;
    return (int64_t)comb_unrank<uint64_t, uint64_t>(r, L, nup, binom);
}

template <typename V>
__global__ __launch_bounds__(kPB) void k_project_stored(GroupDev g, int nup, int kind, const uint32_t *__restrict__ reps,
                                                        const uint32_t *__restrict__ orbit, const uint64_t *binom64,
                                                        int64_t ns, const V *__restrict__ psi, V *__restrict__ out) {
    using T = VTraits<V>;
    __shared__ uint32_t s_binom[33 * 65];
    for (int i = threadIdx.x; i < 33 * 65; i += kPB) s_binom[i] = (uint32_t)binom64[i];
    __syncthreads();
    const uint32_t mask = g.L >= 32 ? 0xFFFFFFFFu : ((1u << g.L) - 1);
    const int order = g.nT * g.npoint;
    for (int64_t c = (int64_t)blockIdx.x * kPB + threadIdx.x; c < ns; c += (int64_t)gridDim.x * kPB) {
        const uint32_t r = reps[c];
        double ar = 0, ai = 0;
        for (int p = 0; p < g.npoint; p++) {
            const uint32_t v = point_apply(g, p, r);
            for (int t = 0; t < g.nT; t++) {
                const uint32_t u = rotl_bits(v, (t * g.a) % g.L, g.L, mask);
                double xr, xi;
                T::load(psi, parent_index(u, nup, kind, s_binom), xr, xi);
                const double lr = g.lam_re[t], li = g.lam_im[t];
                const double pr = g.pch_re[p], pi = g.pch_im[p];
                const double cr = lr * pr - li * pi, ci = lr * pi + li * pr;
                ar += cr * xr - ci * xi;
                ai += cr * xi + ci * xr;
            }
        }
        const double f = sqrt((double)orbit[c]) / (double)order;
        T::store(out, c, ar * f, ai * f);
    }
}

// spin inversion only: |c> = (|s> + z |~s>) / sqrt(2), s with bit L-1 clear;
// without symmetry (zchar = 0): out[c] = psi[state c]
template <typename V>
__global__ void k_project_implicit(int L, int nup, int kind, int zchar, uint64_t cfull, int64_t ns,
                                   const uint64_t *binom, const V *__restrict__ psi, V *__restrict__ out) {
    using T = VTraits<V>;
    const uint64_t mask = L >= 64 ? ~0ULL : ((1ULL << L) - 1);
    const double h = 0.70710678118654752440;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ns; c += (int64_t)gridDim.x * blockDim.x) {
        double ar, ai;
        T::load(psi, own_to_parent(c, L, nup, kind, binom), ar, ai);
        if (zchar == 0) {
            T::store(out, c, ar, ai);
            continue;
        }
        int64_t partner;
        if (kind == 1 || nup < 0) {
            const uint64_t st = nup >= 0 ? comb_unrank<uint64_t, uint64_t>(c, L, nup, binom) : (uint64_t)c;
            partner = (int64_t)(~st & mask);
        } else {
            partner = (int64_t)(cfull - 1 - (uint64_t)c);
        }
        double br, bi;
        T::load(psi, partner, br, bi);
        T::store(out, c, h * (ar + zchar * br), h * (ai + zchar * bi));
    }
}

template <typename V>
__global__ void k_expand(int mode, int L, int nup, int kind, int zchar, uint64_t cfull, int64_t ns, int64_t npar,
                         int64_t ldp, const uint64_t *binom, const uint32_t *__restrict__ ptab, int colbits,
                         GroupDev g, int track, const double *__restrict__ inv_nrm, const V *__restrict__ x,
                         V *__restrict__ parent, int ncols, int accumulate) {
    using T = VTraits<V>;
    const uint64_t mask = L >= 64 ? ~0ULL : ((1ULL << L) - 1);
    const uint32_t colmask = colbits >= 32 ? 0xFFFFFFFFu : ((1u << colbits) - 1);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < npar;
         r += (int64_t)gridDim.x * blockDim.x) {
        // P[r, c] = (ar, ai) for the one column c the parent state belongs to
        int64_t c = -1;
        double ar = 0, ai = 0;
        if (mode == MODE_FULL) {
            c = r;
            ar = 1.0;
        } else if (mode == MODE_ZONLY) {
            ar = 0.70710678118654752440;
            if (nup >= 0) {
                c = r < ns ? r : (int64_t)(cfull - 1 - (uint64_t)r);
                if (r >= ns) ar *= zchar;
            } else {
                const bool top = (r >> (L - 1)) & 1;
                c = top ? (int64_t)(~(uint64_t)r & mask) : r;
                if (top) ar *= zchar;
            }
        } else {
            const uint32_t e = ptab[r];
            if (e != 0xFFFFFFFFu) {
                c = e & colmask;
                double cr = 1.0, ci = 0.0;
                if (track) {
                    const int elem = (int)(e >> colbits);
                    const int bt = elem / g.npoint, bp = elem - bt * g.npoint;
                    const double lr = g.lam_re[bt], li = g.lam_im[bt];
                    const double pr = g.pch_re[bp], pi = g.pch_im[bp];
                    cr = lr * pr - li * pi;
                    ci = lr * pi + li * pr;
                }
                const double f = inv_nrm[c];
                ar = f * cr;
                ai = f * ci;
            }
        }
        const int64_t pr_idx = own_to_parent(r, L, nup, kind, binom);
        for (int j = 0; j < ncols; j++) {
            double vr = 0, vi = 0;
            if (c >= 0) {
                double xr, xi;
                T::load(x, (int64_t)j * ns + c, xr, xi);
                vr = ar * xr - ai * xi;
                vi = ar * xi + ai * xr;
            }
            if (accumulate) {
                double pr, pi;
                T::load(parent, (int64_t)j * ldp + pr_idx, pr, pi);
                vr += pr;
                vi += pi;
            }
            T::store(parent, (int64_t)j * ldp + pr_idx, vr, vi);
        }
    }
}

template <typename F>
void by_dtype(int dtype, F &&f) {
    switch (dtype) {
    case QS_F64: f((double *)nullptr); break;
    case QS_C128: f((cplx_d *)nullptr); break;
    case QS_F32: f((float *)nullptr); break;
    case QS_C64: f((cplx_f *)nullptr); break;
    default: fail(QS_EARG, "unknown dtype");
    }
}

int pgrid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

bool tracks(const GroupHost &g) {
    for (int p = 0; p < g.npoint; p++)
        if (g.pchar[p] != cd(1, 0)) return true;
    for (int t = 0; t < g.nT; t++)
        if (g.lam[t] != cd(1, 0)) return true;
    return false;
}

}  // namespace

int64_t parent_length(const BasisHost &b, int kind) {
    if (b.wide) fail(QS_EARG, "projection to and from the parent is not supported for a wide basis");
    if (kind == 1) {
        if (b.L > 40) fail(QS_EARG, "a state-indexed parent vector needs L <= 40");
        return (int64_t)1 << b.L;
    }
    return b.npar;
}

void basis_project(const BasisHost &b, int dtype, int kind, const void *parent, void *out, hipStream_t s) {
    if (b.group.general) fail(QS_EARG, "projection to and from the parent is not supported for general symmetry maps");
    if (b.is_complex && !dtype_complex(dtype)) fail(QS_EARG, "a complex sector needs a complex vector dtype");
    if (kind != 0 && kind != 1) fail(QS_EARG, "parent kind must be 0 (own parent) or 1 (all states)");
    const size_t esz = dtype_size(dtype);
    if (b.mode == MODE_FULL && (kind == 0 || b.nup < 0)) {
        QS_HIP(hipMemcpyAsync(out, parent, (size_t)b.Ns * esz, hipMemcpyDeviceToDevice, s));
        return;
    }
    const uint64_t cfull = b.nup >= 0 ? (uint64_t)binom(b.L, b.nup) : 0;
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        if (b.mode != MODE_STORED) {
            hipLaunchKernelGGL((k_project_implicit<V>), dim3(pgrid(b.Ns)), dim3(256), 0, s, b.L, b.nup, kind,
                               b.mode == MODE_ZONLY ? b.zchar : 0, cfull, b.Ns, b.binom.get(), (const V *)parent,
                               (V *)out);
        } else {
            if (b.use64) fail(QS_EARG, "sector projection supports L <= 32");
            hipLaunchKernelGGL((k_project_stored<V>), dim3(pgrid(b.Ns)), dim3(kPB), 0, s, group_to_dev(b.group), b.nup,
                               kind, b.reps32.get(), b.orbit.get(), b.binom.get(), b.Ns, (const V *)parent, (V *)out);
        }
    });
    QS_HIP(hipGetLastError());
}

void basis_expand(const BasisHost &b, int dtype, int kind, const void *x, void *parent, int ncols, bool accumulate,
                  hipStream_t s) {
    if (b.group.general) fail(QS_EARG, "projection to and from the parent is not supported for general symmetry maps");
    if (b.is_complex && !dtype_complex(dtype)) fail(QS_EARG, "a complex sector needs a complex vector dtype");
    if (kind != 0 && kind != 1) fail(QS_EARG, "parent kind must be 0 (own parent) or 1 (all states)");
    if (b.mode == MODE_STORED && !b.ptab.get())
        fail(QS_EARG, "sector expansion needs the parent lookup table (L <= 32 and enough device memory)");
    const uint64_t cfull = b.nup >= 0 ? (uint64_t)binom(b.L, b.nup) : 0;
    const int64_t ldp = parent_length(b, kind);
    const size_t esz = dtype_size(dtype);
    if (!accumulate && kind == 1 && b.nup >= 0)
        // states outside the particle sector are not written below
        QS_HIP(hipMemsetAsync(parent, 0, (size_t)ldp * ncols * esz, s));
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_expand<V>), dim3(pgrid(b.npar)), dim3(256), 0, s, b.mode, b.L, b.nup, kind, b.zchar,
                           cfull, b.Ns, b.npar, ldp, b.binom.get(), b.ptab.get(), b.ptab_colbits,
                           group_to_dev(b.group), (int)tracks(b.group), b.inv_nrm.get(), (const V *)x, (V *)parent,
                           ncols, (int)accumulate);
    });
    QS_HIP(hipGetLastError());
}

}  // namespace qs
