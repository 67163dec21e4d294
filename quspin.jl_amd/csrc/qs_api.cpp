// qs_api.cpp -- the extern "C" boundary of libqsgpu (declared in include/qsgpu.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "qs_internal.h"

struct qs_basis {
    qs::BasisHost b;
};
struct qs_op {
    qs::OpHost o;
};

namespace qs {

const std::string &last_error();

OpHost::~OpHost() {
    xs_bufs.clear();   // hipFree: waits for the products that read them
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

char *OpHost::xs_for(hipStream_t s, size_t bytes) const {
    std::lock_guard<std::mutex> g(xs_mutex);
    for (auto &e : xs_bufs) {
        if (e->s != s) continue;
        if (e->buf.n < bytes) e->buf.alloc(bytes);
        return e->buf.get();
    }
    auto e = std::make_unique<XsBuf>();
    e->s = s;
    e->buf.alloc(bytes);
    char *p = e->buf.get();
    xs_bufs.push_back(std::move(e));
    return p;
}

namespace {

template <typename F>
int guard(F &&f) {
    try {
        f();
        return QS_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return QS_EOOM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return QS_EDEVICE;
    }
}

void check_ptr(const void *p, const char *name) {
    if (!p) fail(QS_EARG, std::string(name) + " must not be NULL");
}

// host-side canonicalisation, used by qs_basis_lookup
uint64_t canon_host(const BasisHost &b, uint64_t u, cd &chi) {
    const GroupHost &g = b.group;
    uint64_t best = ~0ULL;
    chi = cd(1, 0);
    if (g.general) {
        for (int e = 0; e < g.order(); e++) {
            const uint64_t w = group_apply(g, e, u);
            if (w < best) {
                best = w;
                chi = g.echar[(size_t)e];
            }
        }
        return best;
    }
    for (int p = 0; p < g.npoint; p++) {
        uint64_t v = apply_point(g, p, u);
        for (int t = 0; t < g.nT; t++) {
            uint64_t w = rotl(v, t * g.a, g.L);
            if (w < best) {
                best = w;
                chi = g.lam[t] * g.pchar[p];
            }
        }
    }
    return best;
}

int64_t rank_host(uint64_t u) {
    int64_t r = 0;
    int k = 1;
    while (u) {
        int p = __builtin_ctzll(u);
        r += binom(p, k);
        k++;
        u &= u - 1;
    }
    return r;
}

uint64_t unrank_host(int64_t i, int L, int w) {
    uint64_t u = 0;
    int p = L - 1;
    for (int k = w; k >= 1; k--) {
        while (binom(p, k) > i) p--;
        u |= 1ULL << p;
        i -= binom(p, k);
        p--;
    }
    return u;
}

// the reference labels and normalizations, computed on first use
std::mutex g_label_mutex;
void ensure_labels(BasisHost &b) {
    std::lock_guard<std::mutex> g(g_label_mutex);
    if (b.labels_ready) return;
    if (b.wide) compute_wide_labels(b);
    else compute_reference_labels(b);
}

void host_reps(const BasisHost &b, std::vector<uint64_t> &out) {
    out.resize((size_t)b.Ns);
    if (b.wide) {
        if (b.L > 64) fail(QS_EARG, "states of more than 64 sites take four words: use qs_basis_states_wide");
        std::vector<uint64_t> w((size_t)4 * b.Ns);
        QS_HIP(hipMemcpy(w.data(), b.wreps.get(), sizeof(uint64_t) * w.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < b.Ns; i++) out[(size_t)i] = w[(size_t)4 * i];
        return;
    }
    if (b.mode == MODE_STORED) {
        if (b.use64) {
            QS_HIP(hipMemcpy(out.data(), b.reps64.get(), sizeof(uint64_t) * b.Ns, hipMemcpyDeviceToHost));
        } else {
            std::vector<uint32_t> tmp((size_t)b.Ns);
            QS_HIP(hipMemcpy(tmp.data(), b.reps32.get(), sizeof(uint32_t) * b.Ns, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < b.Ns; i++) out[i] = tmp[i];
        }
        return;
    }
    if (b.nup < 0) {
        for (int64_t i = 0; i < b.Ns; i++) out[i] = (uint64_t)i;
        return;
    }
    // walk the combinadic order with Gosper's successor
    uint64_t s = unrank_host(0, b.L, b.nup);
    for (int64_t i = 0; i < b.Ns; i++) {
        out[i] = s;
        if (i + 1 == b.Ns || s == 0) break;
        uint64_t lowest = s & (~s + 1), nx = s + lowest;
        s = nx | ((nx ^ s) >> (__builtin_ctzll(lowest) + 2));
    }
}

bool structural_hermitian(const std::vector<Coupling> &cs) {
    // operators.jl:2979-3007
    std::map<std::pair<std::string, std::vector<int>>, cd> coef;
    for (auto &c : cs) coef[{c.op, c.sites}] += c.coef;
    for (auto &kv : coef) {
        std::string adj = kv.first.first;
        for (char &ch : adj) {
            if (ch == '+') ch = '-';
            else if (ch == '-') ch = '+';
            else if (ch != 'I' && ch != 'x' && ch != 'y' && ch != 'z' && ch != 'n' && ch != '|') return false;
        }
        auto it = coef.find({adj, kv.first.second});
        if (it == coef.end()) return false;
        cd want = std::conj(kv.second);
        double d = std::abs(it->second - want);
        double tol = std::max(3e-12, 3e-12 * std::max(std::abs(it->second), std::abs(want)));
        if (d > tol) return false;
    }
    return true;
}

bool pcon_ok(const std::vector<Coupling> &cs) {
    // discrete_basis.jl:2333-2339
    for (auto &c : cs) {
        int change = 0;
        for (char ch : c.op) {
            if (ch == 'x' || ch == 'y') return false;
            if (ch == '+') change++;
            if (ch == '-') change--;
        }
        if (change != 0) return false;
    }
    return true;
}

// adjoint of a compiled coupling (same flip, shifted condition, conj amplitude)
CTerm adjoint_term(const CTerm &t) {
    CTerm a = t;
    a.cval = t.cval ^ (t.flip & t.cmask);
    a.A = std::conj(t.A);
    if (__builtin_popcountll(t.flip & t.smask) & 1) a.A = -a.A;
    return a;
}

void upload_terms(const std::vector<CTerm> &src, DevBuf<TermDev> &dst) {
    std::vector<TermDev> h(std::max<size_t>(1, src.size()));
    for (size_t i = 0; i < src.size(); i++) {
        h[i].cmask = src[i].cmask;
        h[i].cval = src[i].cval;
        h[i].smask = src[i].smask;
        h[i].are = src[i].A.real();
        h[i].aim = src[i].A.imag();
    }
    dst.alloc(h.size());
    QS_HIP(hipMemcpy(dst.get(), h.data(), sizeof(TermDev) * h.size(), hipMemcpyHostToDevice));
}

void upload_groups(const std::vector<GroupOfTerms> &src, DevBuf<GroupOfTerms> &dst) {
    std::vector<GroupOfTerms> h(src);
    if (h.empty()) h.push_back(GroupOfTerms{0, 0, 0, -1, -1});
    dst.alloc(h.size());
    QS_HIP(hipMemcpy(dst.get(), h.data(), sizeof(GroupOfTerms) * h.size(), hipMemcpyHostToDevice));
}

template <typename T>
void upload_vec(const std::vector<T> &src, DevBuf<T> &dst) {
    std::vector<T> h(src);
    if (h.empty()) h.push_back(T{});
    dst.alloc(h.size());
    QS_HIP(hipMemcpy(dst.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

void upload_fast(const FastTerms &f, DevBuf<HopDev> &hops, DevBuf<DiagClass> &dcls, DevBuf<TermDev> &drest) {
    upload_vec(f.hops, hops);
    upload_vec(f.dcls, dcls);
    upload_terms(f.drest, drest);
}

// Device term lists of an operator from its group-averaged couplings.
void setup_terms(OpHost &o, const std::vector<CTerm> &Havg, const double *diag) {
    const BasisHost &bb = *o.basis;
    // pull-form kernels: plain products apply the adjoint couplings
    std::vector<CTerm> Hadj;
    for (auto &x : Havg) Hadj.push_back(adjoint_term(x));
    split_terms(Hadj, o.diag_terms, o.off_terms, o.groups);
    split_terms(Havg, o.diag_terms_adj, o.off_terms_adj, o.groups_adj);
    upload_terms(o.diag_terms, o.d_diag);
    upload_terms(o.off_terms, o.d_off);
    upload_groups(o.groups, o.d_groups);
    upload_terms(o.diag_terms_adj, o.d_diag_adj);
    upload_terms(o.off_terms_adj, o.d_off_adj);
    upload_groups(o.groups_adj, o.d_groups_adj);
    o.fast = compile_fast(o.diag_terms, o.off_terms, o.groups, bb.L, bb.nup);
    o.fast_adj = compile_fast(o.diag_terms_adj, o.off_terms_adj, o.groups_adj, bb.L, bb.nup);
    upload_fast(o.fast, o.d_hops, o.d_dcls, o.d_drest);
    upload_fast(o.fast_adj, o.d_hops_adj, o.d_dcls_adj, o.d_drest_adj);
    if (bb.group.general && !(hop_eligible(o, false) && hop_eligible(o, true)))
        fail(QS_EARG, "a basis of general symmetry maps takes two-bit hop couplings (the parent-table kernels)");
    if (o.has_user_diag) {
        o.d_user_diag.alloc((size_t)2 * bb.Ns);
        QS_HIP(hipMemcpy(o.d_user_diag.get(), diag, sizeof(double) * 2 * bb.Ns, hipMemcpyHostToDevice));
    }
}

hipStream_t ensure_stream(OpHost &o) {
    if (o.stream) return o.stream;
    if (!o.own_stream) QS_HIP(hipStreamCreate(&o.own_stream));
    o.stream = o.own_stream;
    return o.stream;
}

void ensure_stage(DevBuf<char> &buf, size_t bytes) {
    if (buf.n < bytes) buf.alloc(bytes);
}

}  // namespace
}  // namespace qs

namespace qs {
// eigsh on stream s (see qs_eigsh); host v0 / vals / vecs
void eigsh_host(OpHost &o, hipStream_t s, int dtype, int k, int which, double tol, int maxiter, int ncv,
                const void *v0, double *vals, void *vecs, int *nconv, int *n_matvec) {
    const int64_t n = o.basis->Ns;
    if (k < 1 || k > n) fail(QS_EARG, "k must lie in 1:Ns");
    if (which < QS_WHICH_SA || which > QS_WHICH_BE) fail(QS_EARG, "which must be SA, LA, SM, LM, or BE");
    if (o.is_complex && !dtype_complex(dtype)) fail(QS_EARG, "a complex operator needs a complex vector dtype");
    StreamScope scope(s);
    const size_t esz = dtype_size(dtype);
    DevBuf<char> dv0, dvec;
    if (v0) {
        dv0.alloc((size_t)n * esz);
        QS_HIP(hipMemcpyAsync(dv0.get(), v0, (size_t)n * esz, hipMemcpyHostToDevice, s));
    }
    if (vecs) dvec.alloc((size_t)n * k * esz);
    EigshResult res;
    if (!o.hermitian) {
        // Arpack.eigs' complex non-symmetric path (operators.jl:3383-3390);
        // k >= Ns - 1 takes the dense eigen decomposition (:3379-3382)
        if (!dtype_complex(dtype)) fail(QS_EARG, "a non-Hermitian operator needs a complex vector dtype");
        if (which == QS_WHICH_BE) fail(QS_EARG, "which=:BE only possible for real symmetric problem");
        if (k >= n - 1) {
            const int nn = (int)n;
            DevBuf<char> E((size_t)nn * nn * 16), Y((size_t)nn * nn * 16);
            std::vector<cd> eh((size_t)nn * nn, cd(0, 0));
            for (int j = 0; j < nn; j++) eh[(size_t)j * nn + j] = 1.0;
            QS_HIP(hipMemcpyAsync(E.get(), eh.data(), eh.size() * 16, hipMemcpyHostToDevice, s));
            for (int j = 0; j < nn; j++)
                launch_matvec(o, QS_C128, E.get() + (size_t)j * nn * 16, Y.get() + (size_t)j * nn * 16, 1, 0, 0, 0,
                              false, false, false, s);
            QS_HIP(hipMemcpyAsync(eh.data(), Y.get(), eh.size() * 16, hipMemcpyDeviceToHost, s));
            QS_HIP(hipStreamSynchronize(s));
            std::vector<cd> A((size_t)nn * nn), w, Vv;
            for (int i = 0; i < nn; i++)
                for (int j = 0; j < nn; j++) A[(size_t)i * nn + j] = eh[(size_t)j * nn + i];
            gen_eig(nn, A, w, Vv);
            std::vector<int> idx(nn);
            for (int i = 0; i < nn; i++) idx[i] = i;
            auto key = [&](int i) {
                switch (which) {
                case QS_WHICH_LA: return -w[i].real();
                case QS_WHICH_SM: return std::abs(w[i]);
                case QS_WHICH_LM: return -std::abs(w[i]);
                default: return w[i].real();
                }
            };
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b2) { return key(a) < key(b2); });
            for (int i = 0; i < k; i++) {
                vals[2 * i] = w[idx[i]].real();
                vals[2 * i + 1] = w[idx[i]].imag();
            }
            if (vecs) {
                std::vector<char> vh((size_t)n * k * esz);
                for (int c = 0; c < k; c++) {
                    double nrm = 0;
                    for (int r = 0; r < nn; r++) nrm += std::norm(Vv[(size_t)r * nn + idx[c]]);
                    nrm = std::sqrt(nrm);
                    for (int r = 0; r < nn; r++) {
                        const cd z = Vv[(size_t)r * nn + idx[c]] / nrm;
                        const size_t off = (size_t)c * n + r;
                        if (dtype == QS_C128) {
                            ((double *)vh.data())[2 * off] = z.real();
                            ((double *)vh.data())[2 * off + 1] = z.imag();
                        } else {
                            ((float *)vh.data())[2 * off] = (float)z.real();
                            ((float *)vh.data())[2 * off + 1] = (float)z.imag();
                        }
                    }
                }
                memcpy(vecs, vh.data(), vh.size());
            }
            if (nconv) *nconv = k;
            if (n_matvec) *n_matvec = nn;
            return;
        }
        eigs_device(o, dtype, k, which, tol, maxiter, ncv, v0 ? dv0.get() : nullptr, res, vecs ? dvec.get() : nullptr);
        for (int i = 0; i < 2 * k; i++) vals[i] = res.vals[i];
        if (vecs) QS_HIP(hipMemcpyAsync(vecs, dvec.get(), (size_t)n * k * esz, hipMemcpyDeviceToHost, s));
        QS_HIP(hipStreamSynchronize(s));
        if (nconv) *nconv = res.nconv;
        if (n_matvec) *n_matvec = res.nmatvec;
        if (res.nconv < k)
            fail(QS_ENOCONV, "eigsh: only " + std::to_string(res.nconv) + " of " + std::to_string(k) +
                                 " Ritz values converged within maxiter");
        return;
    }
    if (k >= n - 1) {
        // operators.jl:3379-3382: dense fallback for tiny problems
        int nn = (int)n;
        int cdt = dtype_complex(dtype) ? QS_C128 : QS_F64;
        DevBuf<char> E((size_t)nn * nn * dtype_size(cdt)), Y((size_t)nn * nn * dtype_size(cdt));
        std::vector<char> eh((size_t)nn * nn * dtype_size(cdt), 0);
        for (int j = 0; j < nn; j++) {
            if (cdt == QS_F64) ((double *)eh.data())[(size_t)j * nn + j] = 1.0;
            else ((double *)eh.data())[2 * ((size_t)j * nn + j)] = 1.0;
        }
        QS_HIP(hipMemcpyAsync(E.get(), eh.data(), eh.size(), hipMemcpyHostToDevice, s));
        for (int j = 0; j < nn; j++)
            launch_matvec(o, cdt, E.get() + (size_t)j * nn * dtype_size(cdt), Y.get() + (size_t)j * nn * dtype_size(cdt),
                          1, 0, 0, 0, false, false, false, s);
        QS_HIP(hipMemcpyAsync(eh.data(), Y.get(), eh.size(), hipMemcpyDeviceToHost, s));
        QS_HIP(hipStreamSynchronize(s));
        std::vector<cd> A((size_t)nn * nn);
        for (int i = 0; i < nn; i++)
            for (int j = 0; j < nn; j++)
                A[(size_t)i * nn + j] = cdt == QS_F64 ? cd(((double *)eh.data())[(size_t)j * nn + i], 0)
                                                      : cd(((double *)eh.data())[2 * ((size_t)j * nn + i)],
                                                           ((double *)eh.data())[2 * ((size_t)j * nn + i) + 1]);
        std::vector<double> w;
        std::vector<cd> Vv;
        herm_eig(nn, A, w, Vv);
        std::vector<int> idx(nn);
        for (int i = 0; i < nn; i++) idx[i] = i;
        auto key = [&](int i) {
            switch (which) {
            case QS_WHICH_LA: return -w[i];
            case QS_WHICH_SM: return std::fabs(w[i]);
            case QS_WHICH_LM: return -std::fabs(w[i]);
            default: return w[i];
            }
        };
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b2) { return key(a) < key(b2); });
        if (which == QS_WHICH_BE) {
            std::vector<int> be;
            for (int i = 0; i < k / 2; i++) be.push_back(i);
            for (int i = 0; i < (k + 1) / 2; i++) be.push_back(nn - 1 - i);
            idx = be;
        }
        for (int i = 0; i < k; i++) vals[i] = w[idx[i]];
        if (vecs) {
            std::vector<char> vh((size_t)n * k * esz);
            for (int c = 0; c < k; c++)
                for (int r = 0; r < nn; r++) {
                    cd z = Vv[(size_t)r * nn + idx[c]];
                    size_t off = (size_t)c * n + r;
                    switch (dtype) {
                    case QS_F64: ((double *)vh.data())[off] = z.real(); break;
                    case QS_F32: ((float *)vh.data())[off] = (float)z.real(); break;
                    case QS_C128: ((double *)vh.data())[2 * off] = z.real(); ((double *)vh.data())[2 * off + 1] = z.imag(); break;
                    case QS_C64: ((float *)vh.data())[2 * off] = (float)z.real(); ((float *)vh.data())[2 * off + 1] = (float)z.imag(); break;
                    }
                }
            memcpy(vecs, vh.data(), vh.size());
        }
        if (nconv) *nconv = k;
        if (n_matvec) *n_matvec = nn;
        return;
    }
    eigsh_device(o, dtype, k, which, tol, maxiter, ncv, v0 ? dv0.get() : nullptr, res, vecs ? dvec.get() : nullptr);
    for (int i = 0; i < k; i++) vals[i] = res.vals[i];
    if (vecs) QS_HIP(hipMemcpyAsync(vecs, dvec.get(), (size_t)n * k * esz, hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    if (nconv) *nconv = res.nconv;
    if (n_matvec) *n_matvec = res.nmatvec;
    if (res.nconv < k) fail(QS_ENOCONV, "eigsh: only " + std::to_string(res.nconv) + " of " + std::to_string(k) +
                                            " Ritz values converged within maxiter");
}
// _krylov_expmv_times on stream s (see qs_expm_multiply); host v / out
void expm_host(OpHost &o, hipStream_t s, int dtype, const void *v, int64_t n_times, const double *scales, double tol,
               int krylov_dim, void *out, int *n_matvec) {
    StreamScope scope(s);
    const int64_t n = o.basis->Ns;
    const size_t esz = dtype_size(dtype);
    std::vector<cd> sc((size_t)n_times);
    for (int64_t j = 0; j < n_times; j++) sc[j] = cd(scales[2 * j], scales[2 * j + 1]);
    DevBuf<char> dv((size_t)n * esz), dout((size_t)n * std::max<int64_t>(1, n_times) * esz);
    QS_HIP(hipMemcpyAsync(dv.get(), v, (size_t)n * esz, hipMemcpyHostToDevice, s));
    int nmv = 0;
    expm_times_device(o, dtype, dv.get(), sc, tol, krylov_dim, dout.get(), nmv);
    QS_HIP(hipMemcpyAsync(out, dout.get(), (size_t)n * n_times * esz, hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    if (n_matvec) *n_matvec = nmv;
}
}  // namespace qs

using namespace qs;

extern "C" {

size_t qs_last_error(char *buf, size_t len) {
    const std::string &e = last_error();
    if (buf && len) {
        size_t n = std::min(len - 1, e.size());
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return e.size();
}

int qs_device_info(char *name, size_t len, int64_t *hbm_bytes, int *n_cu) {
    return guard([&] {
        int dev = 0;
        QS_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        QS_HIP(hipGetDeviceProperties(&p, dev));
        if (name && len) {
            std::string s = std::string(p.name) + " (" + p.gcnArchName + ")";
            size_t n = std::min(len - 1, s.size());
            memcpy(name, s.data(), n);
            name[n] = 0;
        }
        if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
        if (n_cu) *n_cu = p.multiProcessorCount;
    });
}

int qs_basis_create(int L, int nup, int pauli, int a, int kblock, int pblock, int zblock, int pzblock, int zAblock,
                    int zBblock, int flags, qs_basis **out) {
    (void)flags;
    return guard([&] {
        check_ptr(out, "out");
        *out = nullptr;
        // basis.jl:652-665
        if (L < 1 || L > 63) fail(QS_EARG, "L must be between 1 and 63");
        if (nup != QS_UNSET && (nup < 0 || nup > L)) fail(QS_EARG, "nup must lie between 0 and L");
        auto h = std::make_unique<qs_basis>();
        BasisHost &b = h->b;
        b.L = L;
        b.nup = nup == QS_UNSET ? -1 : nup;
        b.pauli = pauli ? 1 : 0;
        b.a = a;
        int blocks[6] = {kblock, pblock, zblock, pzblock, zAblock, zBblock};
        memcpy(b.blocks, blocks, sizeof blocks);
        if (b.nup < 0 && L > 40) fail(QS_EARG, "a basis without nup is limited to L <= 40 here");
        build_group(b);
        build_basis_device(b);
        *out = h.release();
    });
}

int qs_basis_create_general(int L, int nup, int pauli, int nblocks, const int32_t *maps, const int32_t *qnumbers,
                            int flags, qs_basis **out) {
    return guard([&] {
        check_ptr(out, "out");
        *out = nullptr;
        if (L < 1 || L > kMaxWideL) fail(QS_EARG, "L must be between 1 and 256");
        const bool wide = (flags & QS_BASIS_EXPLICIT) || L > 63 || (L > 32 && nblocks > 0);
        if (wide && nup == QS_UNSET)
            fail(QS_EARG, "wide spin bases require nup, Nup, or m to avoid enumerating 2^L states");
        if (nup != QS_UNSET && (nup < 0 || nup > L)) fail(QS_EARG, "nup must lie between 0 and L");
        if (nblocks < 0) fail(QS_EARG, "nblocks must be non-negative");
        if (nblocks > 0) {
            check_ptr(maps, "maps");
            check_ptr(qnumbers, "qnumbers");
        }
        auto h = std::make_unique<qs_basis>();
        BasisHost &b = h->b;
        b.L = L;
        b.nup = nup == QS_UNSET ? -1 : nup;
        b.pauli = pauli ? 1 : 0;
        b.a = 1;
        for (int i = 0; i < 6; i++) b.blocks[i] = QS_UNSET;
        if (b.nup < 0 && L > 40) fail(QS_EARG, "a basis without nup is limited to L <= 40 here");
        for (int bi = 0; bi < nblocks; bi++) {
            // basis.jl:282-304: QuSpin's zero-based labels when any entry is
            // zero or negative (-v: flipped, to site v), else one-based
            const int32_t *v = maps + (size_t)bi * L;
            bool zero_based = false;
            for (int i = 0; i < L; i++) zero_based |= v[i] <= 0;
            GenMap m;
            m.perm.assign((size_t)L, -1);
            std::vector<char> hit((size_t)L + 1, 0);
            bool ok = true;
            for (int i = 0; i < L; i++) {
                const int64_t target = zero_based ? (v[i] >= 0 ? (int64_t)v[i] + 1 : -(int64_t)v[i])
                                                  : (v[i] < 0 ? -(int64_t)v[i] : (int64_t)v[i]);
                if (v[i] < 0) {
                    m.flipw.set(i);
                    if (i < 64) m.flip |= 1ULL << i;
                }
                if (target < 1 || target > L || hit[(size_t)target]) {
                    ok = false;
                    break;
                }
                hit[(size_t)target] = 1;
                m.perm[(size_t)i] = (int)target - 1;
            }
            if (!ok) fail(QS_EARG, "a general symmetry map must be a site permutation");
            // basis.jl:306-323: lcm of the cycle lengths, doubled for cycles with an odd number of flips
            std::vector<char> seen((size_t)L, 0);
            int64_t order = 1;
            for (int st = 0; st < L; st++) {
                if (seen[(size_t)st]) continue;
                int cur = st, len = 0, par = 0;
                while (!seen[(size_t)cur]) {
                    seen[(size_t)cur] = 1;
                    len++;
                    par ^= (int)m.flipw.bit(cur);
                    cur = m.perm[(size_t)cur];
                }
                const int64_t c = par ? 2 * len : len;
                order = order / std::gcd(order, c) * c;
                if (order > kMaxGeneral) fail(QS_EARG, "symmetry group too large");
            }
            m.order = (int)order;
            m.q = qnumbers[bi];
            b.gmaps.push_back(m);
        }
        if (wide) {
            build_wide_basis(b);
        } else {
            build_group_general(b);
            build_basis_device(b);
        }
        *out = h.release();
    });
}

int qs_basis_destroy(qs_basis *b) {
    delete b;
    return QS_OK;
}

int qs_basis_size(const qs_basis *b, int64_t *ns) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(ns, "ns");
        *ns = b->b.Ns;
    });
}

int qs_basis_reps(const qs_basis *b, uint64_t *out) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(out, "out");
        std::vector<uint64_t> r;
        host_reps(b->b, r);
        memcpy(out, r.data(), sizeof(uint64_t) * r.size());
    });
}

int qs_basis_states(const qs_basis *b, uint64_t *out) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(out, "out");
        BasisHost &bb = const_cast<BasisHost &>(b->b);
        if (bb.mode != MODE_STORED) {
            std::vector<uint64_t> r;
            host_reps(bb, r);
            memcpy(out, r.data(), sizeof(uint64_t) * r.size());
            return;
        }
        if (bb.wide && bb.L > 64) fail(QS_EARG, "states of more than 64 sites take four words: use qs_basis_states_wide");
        ensure_labels(bb);
        memcpy(out, bb.labels.data(), sizeof(uint64_t) * bb.labels.size());
    });
}

int qs_basis_states_wide(const qs_basis *b, int which, uint64_t *out) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(out, "out");
        if (which != 0 && which != 1) fail(QS_EARG, "which must be 0 (states) or 1 (orbit minima)");
        BasisHost &bb = const_cast<BasisHost &>(b->b);
        const int64_t n = bb.Ns;
        if (bb.wide) {
            if (which == 1) {
                QS_HIP(hipMemcpy(out, bb.wreps.get(), sizeof(uint64_t) * 4 * n, hipMemcpyDeviceToHost));
                return;
            }
            ensure_labels(bb);
            for (int64_t i = 0; i < n; i++)
                for (int k = 0; k < 4; k++) out[4 * i + k] = bb.wlabels[(size_t)i].w[k];
            return;
        }
        std::vector<uint64_t> r;
        if (which == 1 || bb.mode != MODE_STORED) {
            host_reps(bb, r);
        } else {
            ensure_labels(bb);
            r = bb.labels;
        }
        for (int64_t i = 0; i < n; i++) {
            out[4 * i] = r[(size_t)i];
            out[4 * i + 1] = out[4 * i + 2] = out[4 * i + 3] = 0;
        }
    });
}

int qs_basis_norms(const qs_basis *b, uint32_t *orbit_size, double *ref_norm) {
    return guard([&] {
        check_ptr(b, "basis");
        BasisHost &bb = const_cast<BasisHost &>(b->b);
        const int64_t n = bb.Ns;
        if (bb.mode == MODE_FULL) {
            if (orbit_size) std::fill(orbit_size, orbit_size + n, 1u);
            if (ref_norm) std::fill(ref_norm, ref_norm + n, 1.0);
            return;
        }
        if (bb.mode == MODE_ZONLY) {
            // column [1, z]/norm([1, z]) (symmetry_basis.jl:90-104), normalization 1/abs2(amp)
            double nrm = std::sqrt(1.0 * 1.0 + 1.0 * 1.0);
            double amp = 1.0 / nrm;
            if (orbit_size) std::fill(orbit_size, orbit_size + n, 2u);
            if (ref_norm) std::fill(ref_norm, ref_norm + n, 1.0 / (amp * amp));
            return;
        }
        if (orbit_size) QS_HIP(hipMemcpy(orbit_size, bb.orbit.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        if (ref_norm) {
            ensure_labels(bb);
            memcpy(ref_norm, bb.ref_norm.data(), sizeof(double) * n);
        }
    });
}

int qs_basis_is_complex(const qs_basis *b, int *is_complex) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(is_complex, "is_complex");
        *is_complex = b->b.is_complex;
    });
}

int qs_basis_lookup(const qs_basis *b, const uint64_t *states, int64_t n, int64_t *index, double *amplitude) {
    return guard([&] {
        check_ptr(b, "basis");
        const BasisHost &bb = b->b;
        if (bb.wide && bb.L > 64) fail(QS_EARG, "states of more than 64 sites do not fit the 64-bit lookup");
        std::vector<uint64_t> reps;
        if (bb.mode == MODE_STORED) host_reps(bb, reps);
        uint64_t mask = full_mask(bb.L);
        for (int64_t i = 0; i < n; i++) {
            uint64_t s = states[i];
            int64_t idx = -1;
            cd amp(0, 0);
            bool valid = (s & ~mask) == 0 && (bb.nup < 0 || __builtin_popcountll(s) == bb.nup);
            if (valid) {
                if (bb.mode == MODE_FULL) {
                    idx = bb.nup >= 0 ? rank_host(s) : (int64_t)s;
                    amp = 1.0;
                } else if (bb.mode == MODE_ZONLY) {
                    uint64_t r = s;
                    cd ch = 1.0;
                    if (s >> (bb.L - 1) & 1) {
                        r = ~s & mask;
                        ch = (double)bb.zchar;
                    }
                    idx = bb.nup >= 0 ? rank_host(r) : (int64_t)r;
                    amp = ch / std::sqrt(2.0);
                } else {
                    cd ch;
                    uint64_t r = canon_host(bb, s, ch);
                    auto it = std::lower_bound(reps.begin(), reps.end(), r);
                    if (it != reps.end() && *it == r) {
                        idx = it - reps.begin();
                        uint32_t d;
                        QS_HIP(hipMemcpy(&d, bb.orbit.get() + idx, sizeof d, hipMemcpyDeviceToHost));
                        amp = ch / std::sqrt((double)d);
                    }
                }
            }
            if (index) index[i] = idx;
            if (amplitude) {
                amplitude[2 * i] = amp.real();
                amplitude[2 * i + 1] = amp.imag();
            }
        }
    });
}

int qs_basis_lookup_wide(const qs_basis *b, const uint64_t *states, int64_t n, int64_t *index, double *amplitude) {
    return guard([&] {
        check_ptr(b, "basis");
        if (n > 0) check_ptr(states, "states");
        const BasisHost &bb = b->b;
        if (!bb.wide) {
            // states of at most 64 sites: the one-word lookup, beyond-word bits invalid
            std::vector<uint64_t> one((size_t)n);
            std::vector<char> bad((size_t)n, 0);
            for (int64_t i = 0; i < n; i++) {
                one[(size_t)i] = states[4 * i];
                bad[(size_t)i] = (states[4 * i + 1] | states[4 * i + 2] | states[4 * i + 3]) != 0;
            }
            std::vector<int64_t> idx((size_t)std::max<int64_t>(n, 1));
            std::vector<double> amp((size_t)2 * std::max<int64_t>(n, 1));
            const int st = qs_basis_lookup(b, one.data(), n, idx.data(), amp.data());
            if (st != QS_OK) {
                char msg[512];
                qs_last_error(msg, sizeof msg);
                fail(st, msg);
            }
            for (int64_t i = 0; i < n; i++) {
                if (index) index[i] = bad[(size_t)i] ? -1 : idx[(size_t)i];
                if (amplitude) {
                    amplitude[2 * i] = bad[(size_t)i] ? 0.0 : amp[2 * (size_t)i];
                    amplitude[2 * i + 1] = bad[(size_t)i] ? 0.0 : amp[2 * (size_t)i + 1];
                }
            }
            return;
        }
        // wide basis: orbit minimum under the group (first element reaching
        // it, its character), column by binary search over the minima
        const GroupHost &g = bb.group;
        std::vector<uint64_t> words((size_t)4 * bb.Ns);
        QS_HIP(hipMemcpy(words.data(), bb.wreps.get(), sizeof(uint64_t) * words.size(), hipMemcpyDeviceToHost));
        std::vector<uint32_t> orb((size_t)bb.Ns);
        QS_HIP(hipMemcpy(orb.data(), bb.orbit.get(), sizeof(uint32_t) * orb.size(), hipMemcpyDeviceToHost));
        auto rep = [&](int64_t c) {
            WState w;
            for (int k = 0; k < 4; k++) w.w[k] = words[(size_t)4 * c + k];
            return w;
        };
        for (int64_t i = 0; i < n; i++) {
            WState u;
            for (int k = 0; k < 4; k++) u.w[k] = states[4 * i + k];
            bool valid = u.popcount() == bb.nup;
            for (int site = bb.L; site < 256 && valid; site++) valid = !u.bit(site);
            int64_t idx = -1;
            cd amp(0, 0);
            if (valid) {
                WState best = u;
                cd ch = g.echar[0];
                for (int e = 1; e < g.order(); e++) {
                    const WState w = wide_group_apply(g, e, u);
                    if (w < best) {
                        best = w;
                        ch = g.echar[(size_t)e];
                    }
                }
                int64_t lo = 0, hi = bb.Ns;
                while (lo < hi) {
                    const int64_t mid = lo + (hi - lo) / 2;
                    if (rep(mid) < best) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < bb.Ns && rep(lo) == best) {
                    idx = lo;
                    amp = ch / std::sqrt((double)orb[(size_t)lo]);
                }
            }
            if (index) index[i] = idx;
            if (amplitude) {
                amplitude[2 * i] = amp.real();
                amplitude[2 * i + 1] = amp.imag();
            }
        }
    });
}

int qs_op_create(const qs_basis *b, int n_terms, const char *const *opstr, const int32_t *n_couplings,
                 const double *coeff, const int32_t *sites, const double *diag, int flags, qs_op **out) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(out, "out");
        *out = nullptr;
        const BasisHost &bb = b->b;
        std::vector<Coupling> cs;
        size_t ci = 0, si = 0;
        bool any_complex = false;
        for (int t = 0; t < n_terms; t++) {
            std::string op(opstr[t]);
            for (int j = 0; j < n_couplings[t]; j++) {
                Coupling c;
                c.op = op;
                c.coef = cd(coeff[2 * ci], coeff[2 * ci + 1]);
                ci++;
                for (size_t q = 0; q < op.size(); q++) c.sites.push_back(sites[si++]);
                if (c.coef.imag() != 0) any_complex = true;
                cs.push_back(c);
            }
            if (op.find('y') != std::string::npos) any_complex = true;
        }
        auto h = std::make_unique<qs_op>();
        OpHost &o = h->o;
        o.basis = &bb;
        // operators.jl:2519-2534: checks in the reference's order
        bool herm = structural_hermitian(cs);
        bool check_herm = flags & QS_CHECK_HERM, check_pcon = flags & QS_CHECK_PCON, check_symm = flags & QS_CHECK_SYMM;
        if (check_herm && !herm) fail(QS_EARG, "operator list is not Hermitian");
        if (check_pcon && bb.nup >= 0 && !pcon_ok(cs)) fail(QS_EARG, "operator list violates the selected particle sector");
        if (bb.wide) {
            // bases of up to 256 sites: four-word term masks, one kernel
            bool diag_complex = false;
            if (diag) {
                o.user_diag.resize((size_t)bb.Ns);
                for (int64_t i = 0; i < bb.Ns; i++) {
                    o.user_diag[i] = cd(diag[2 * i], diag[2 * i + 1]);
                    if (diag[2 * i + 1] != 0) diag_complex = true;
                }
                o.has_user_diag = true;
                if (check_herm && diag_complex) fail(QS_EARG, "a Hermitian operator requires a real diagonal");
            }
            o.is_complex = (any_complex || bb.is_complex || diag_complex || (flags & 8)) ? 1 : 0;
            setup_wide_terms(o, cs, check_symm);
            o.hermitian = check_herm ? 1 : ((herm && !diag_complex) ? 1 : 0);
            auto upload_diag = [&](OpHost &t) {
                if (!t.has_user_diag) return;
                t.d_user_diag.alloc((size_t)2 * bb.Ns);
                QS_HIP(hipMemcpy(t.d_user_diag.get(), diag, sizeof(double) * 2 * bb.Ns, hipMemcpyHostToDevice));
            };
            upload_diag(o);
            if (bb.group.order() == 1) {
                // a basis without symmetry: the structurally diagonal terms
                // apart, as on 64-bit states (operators.jl:2372-2429)
                std::vector<Coupling> dcs;
                for (auto &c : cs)
                    if (c.op.find_first_not_of("Izn|") == std::string::npos) dcs.push_back(c);
                o.extracted = std::make_unique<OpHost>();
                OpHost &e = *o.extracted;
                e.basis = &bb;
                e.is_complex = o.is_complex;
                e.has_user_diag = o.has_user_diag;
                e.user_diag = o.user_diag;
                setup_wide_terms(e, dcs, false);
                e.is_complex = o.is_complex;
                upload_diag(e);
            }
            ensure_stream(o);
            *out = h.release();
            return;
        }
        std::vector<CTerm> compiled = compile_couplings(cs, bb.L, bb.pauli);
        std::vector<CTerm> H = consolidate_cterms(compiled);
        if (check_symm && bb.mode != MODE_FULL) {
            std::vector<CTerm> avg = symmetrize_cterms(H, bb);
            std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>, cd> a1, a2;
            for (auto &x : H) a1[std::make_tuple(x.flip, x.cmask, x.cval, x.smask)] = x.A;
            for (auto &x : avg) a2[std::make_tuple(x.flip, x.cmask, x.cval, x.smask)] = x.A;
            bool same = a1.size() == a2.size();
            for (auto &kv : a1) {
                auto it = a2.find(kv.first);
                if (it == a2.end() || std::abs(it->second - kv.second) > 1e-12 * std::max(1.0, std::abs(kv.second)))
                    same = false;
            }
            if (!same) fail(QS_EARG, "operator list violates the selected symmetry sector");
        }
        std::vector<CTerm> Havg = symmetrize_cterms(H, bb);
        bool diag_complex = false;
        if (diag) {
            o.user_diag.resize((size_t)bb.Ns);
            for (int64_t i = 0; i < bb.Ns; i++) {
                o.user_diag[i] = cd(diag[2 * i], diag[2 * i + 1]);
                if (diag[2 * i + 1] != 0) diag_complex = true;
            }
            o.has_user_diag = true;
            if (check_herm && diag_complex) fail(QS_EARG, "a Hermitian operator requires a real diagonal");
        }
        o.is_complex = (any_complex || bb.is_complex || diag_complex || (flags & 8)) ? 1 : 0;
        for (auto &x : Havg)
            if (x.A.imag() != 0) o.is_complex = 1;
        o.hermitian = check_herm ? 1 : ((herm && !diag_complex) ? 1 : 0);
        setup_terms(o, Havg, diag);
        if (bb.mode == MODE_FULL) {
            // operators.jl:2372-2429 extracts the structurally diagonal terms
            // of a basis without symmetry; mul! scales them (and the user
            // diagonal) by alpha (:3097-3117) but conjugates alpha together
            // with the other couplings' amplitudes (:2581)
            std::vector<Coupling> dcs;
            for (auto &c : cs)
                if (c.op.find_first_not_of("Izn|") == std::string::npos) dcs.push_back(c);
            o.extracted = std::make_unique<OpHost>();
            OpHost &e = *o.extracted;
            e.basis = &bb;
            e.is_complex = o.is_complex;
            e.has_user_diag = o.has_user_diag;
            e.user_diag = o.user_diag;
            setup_terms(e, consolidate_cterms(compile_couplings(dcs, bb.L, bb.pauli)), diag);
        }
        ensure_stream(o);
        *out = h.release();
    });
}

int qs_block_gram(int dtype, const void *V, int64_t n, int m, const void *W, int p, int on_mfma, double *G) {
    return guard([&] {
        if (!V || !W || !G) fail(QS_EARG, "V, W and G must not be NULL");
        if (n < 1 || m < 1 || p < 1) fail(QS_EARG, "n, m and p must be positive");
        const size_t esz = dtype_size(dtype);
        hipStream_t s;
        QS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};
        VecWork wk(s);
        // bit 1 of the flags: V and W are device arrays
        const bool dev = (on_mfma & 2) != 0;
        SBuf<char> dv(dev ? 0 : (size_t)n * m * esz, s), dw(dev ? 0 : (size_t)n * p * esz, s);
        if (!dev) {
            QS_HIP(hipMemcpyAsync(dv.get(), V, (size_t)n * m * esz, hipMemcpyHostToDevice, s));
            QS_HIP(hipMemcpyAsync(dw.get(), W, (size_t)n * p * esz, hipMemcpyHostToDevice, s));
        }
        const char *pv = dev ? (const char *)V : dv.get(), *pw = dev ? (const char *)W : dw.get();
        std::vector<cd> g;
        if (on_mfma & 1) {
            gram_block(dtype, pv, n, m, pw, p, wk, s, g);
        } else {
            g.resize((size_t)m * p);
            for (int c = 0; c < p; c++) {
                std::vector<cd> h;
                gram_dots(dtype, pv, n, m, pw + (size_t)c * n * esz, wk, s, h);
                for (int i = 0; i < m; i++) g[(size_t)c * m + i] = h[i];
            }
        }
        QS_HIP(hipStreamSynchronize(s));
        for (size_t e = 0; e < g.size(); e++) {
            G[2 * e] = g[e].real();
            G[2 * e + 1] = g[e].imag();
        }
    });
}

int qs_op_destroy(qs_op *op) {
    delete op;
    return QS_OK;
}

int qs_op_info(const qs_op *op, int *is_complex, int *is_hermitian, int64_t *ns) {
    return guard([&] {
        check_ptr(op, "op");
        if (is_complex) *is_complex = op->o.is_complex;
        if (is_hermitian) *is_hermitian = op->o.hermitian;
        if (ns) *ns = op->o.basis->Ns;
    });
}

int qs_op_set_diagonal(qs_op *op, const double *diag) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(diag, "diag");
        int64_t n = op->o.basis->Ns;
        for (OpHost *o : {&op->o, op->o.extracted.get()}) {
            if (!o) continue;
            o->user_diag.resize((size_t)n);
            for (int64_t i = 0; i < n; i++) o->user_diag[i] = cd(diag[2 * i], diag[2 * i + 1]);
            o->has_user_diag = true;
            o->d_user_diag.alloc((size_t)2 * n);
            QS_HIP(hipMemcpy(o->d_user_diag.get(), diag, sizeof(double) * 2 * n, hipMemcpyHostToDevice));
        }
    });
}

int qs_op_set_stream(qs_op *op, void *stream) {
    return guard([&] {
        check_ptr(op, "op");
        op->o.stream = stream ? (hipStream_t)stream : nullptr;
        ensure_stream(op->o);
    });
}

int qs_op_stream(const qs_op *op, void **stream) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(stream, "stream");
        *stream = (void *)op->o.stream;
    });
}

int qs_op_set_variant(qs_op *op, int lanes_per_row) {
    return guard([&] {
        check_ptr(op, "op");
        if (op->o.basis->wide && lanes_per_row != 0)
            fail(QS_EARG, "a wide basis has one matvec kernel (variant 0)");
        if (lanes_per_row != 0 && lanes_per_row != 1 && lanes_per_row != 2 && lanes_per_row != 4 &&
            lanes_per_row != 8 && (lanes_per_row < QS_VARIANT_SPLIT || lanes_per_row > QS_VARIANT_SPLIT_MAX) &&
            (lanes_per_row < QS_VARIANT_HOP || lanes_per_row > QS_VARIANT_HOP + 3) && lanes_per_row != QS_VARIANT_SELL)
            fail(QS_EARG, "variant must be 0, 1, 2, 4, 8, QS_VARIANT_HOP + 0..3, QS_VARIANT_SPLIT + 0..7 or "
                          "QS_VARIANT_SELL");
        if (lanes_per_row == QS_VARIANT_SELL && !op->o.sell) {
            // assembled by the first plain product, on that product's stream
            if (!hop_eligible(op->o, false))
                fail(QS_EARG, "assembly needs two-bit hop couplings, L <= 32 and a parent lookup table");
            op->o.want_sell = true;
        }
        if (lanes_per_row >= QS_VARIANT_SPLIT && lanes_per_row <= QS_VARIANT_SPLIT_MAX &&
            !(split_eligible(op->o, false) && split_eligible(op->o, true)))
            fail(QS_EARG, "the split matvec variant needs two-bit hop couplings on a basis with a particle number "
                          "and no symmetry or spin inversion only, L <= 32");
        if (lanes_per_row >= QS_VARIANT_SPLIT + QS_SPLIT_LEAN && lanes_per_row <= QS_VARIANT_SPLIT_MAX &&
            !(split_lean_ok(op->o, false) && split_lean_ok(op->o, true)))
            fail(QS_EARG, "the lean split kernel does not fit this coupling list (at most 32 hop groups inside "
                          "the low word, 16 across the split and 64 in the high word)");
        if (lanes_per_row >= QS_VARIANT_HOP && lanes_per_row <= QS_VARIANT_HOP + 3 &&
            !(hop_eligible(op->o, false) && hop_eligible(op->o, true)))
            fail(QS_EARG, "the hop matvec variant needs two-bit hop couplings, L <= 32 and a parent lookup table");
        op->o.lanes_per_row = lanes_per_row;
    });
}

int qs_op_tune(qs_op *op, int dtype, int repeats, int *lanes_per_row, double *best_ms) {
    return guard([&] {
        check_ptr(op, "op");
        OpHost &o = op->o;
        if (o.is_complex && !dtype_complex(dtype)) fail(QS_EARG, "a complex operator needs a complex vector dtype");
        hipStream_t s = ensure_stream(o);
        const int64_t n = o.basis->Ns;
        DevBuf<char> x((size_t)n * dtype_size(dtype)), y((size_t)n * dtype_size(dtype));
        vec_random(dtype, x.get(), n, 0x7a11ULL, s);
        hipEvent_t e0, e1;
        QS_HIP(hipEventCreate(&e0));
        QS_HIP(hipEventCreate(&e1));
        int best = 0;
        float best_t = 1e30f;
        if (repeats < 1) repeats = 3;
        std::vector<int> cands;
        if (hop_eligible(o, false))
            cands.insert(cands.end(), {QS_VARIANT_HOP, QS_VARIANT_HOP + 1, QS_VARIANT_HOP + 2, QS_VARIANT_HOP + 3});
        if (split_eligible(o, false) && split_eligible(o, true))
            for (int k = QS_VARIANT_SPLIT; k <= QS_VARIANT_SPLIT_MAX; k++)
                if (k < QS_VARIANT_SPLIT + QS_SPLIT_LEAN || (split_lean_ok(o, false) && split_lean_ok(o, true)))
                    cands.push_back(k);
        if (o.sell) cands.push_back(QS_VARIANT_SELL);
        if (o.basis->wide) cands.assign(1, 0);
        else if (o.basis->mode == MODE_STORED) cands.insert(cands.end(), {1, 2, 4, 8});
        else cands.push_back(1);
        for (int lpr : cands) {
            o.lanes_per_row = lpr;
            launch_matvec(o, dtype, x.get(), y.get(), 1, 0, 0, 0, false, false, false, s);
            QS_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < repeats; r++)
                launch_matvec(o, dtype, x.get(), y.get(), 1, 0, 0, 0, false, false, false, s);
            QS_HIP(hipEventRecord(e1, s));
            QS_HIP(hipEventSynchronize(e1));
            float ms = 0;
            QS_HIP(hipEventElapsedTime(&ms, e0, e1));
            ms /= repeats;
            if (ms < best_t) {
                best_t = ms;
                best = lpr;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        o.lanes_per_row = best;
        if (lanes_per_row) *lanes_per_row = best;
        if (best_ms) *best_ms = best_t;
    });
}

int qs_op_set_graphs(qs_op *op, int enable) {
    return guard([&] {
        check_ptr(op, "op");
        op->o.graphs = enable ? 1 : 0;
    });
}

int qs_op_assemble(qs_op *op) {
    return guard([&] {
        check_ptr(op, "op");
        OpHost &o = op->o;
        if (o.basis->wide) fail(QS_EARG, "assembly is not available for a wide basis");
        hipStream_t s = ensure_stream(o);
        sell_assemble(o, s);
    });
}

int qs_op_csr(const qs_op *op, int64_t *nnz, int64_t *rowptr, int64_t *col, double *val) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(nnz, "nnz");
        const OpHost &o = op->o;
        if (!o.sell) fail(QS_EARG, "the operator is not assembled (qs_op_assemble)");
        std::vector<int64_t> rp, c;
        std::vector<double> v;
        sell_to_csr(o, rp, c, v);
        const int64_t have = *nnz;
        *nnz = (int64_t)c.size();
        if (!rowptr && !col && !val) return;   // size query
        if (have < (int64_t)c.size()) fail(QS_EDIM, "nnz smaller than the assembled matrix");
        if (rowptr) std::copy(rp.begin(), rp.end(), rowptr);
        if (col) std::copy(c.begin(), c.end(), col);
        if (val) std::copy(v.begin(), v.end(), val);
    });
}

int qs_op_matvec(qs_op *op, int dtype, const void *x, void *y, int64_t ncols, double alpha_re, double alpha_im,
                 double beta_re, double beta_im, int transposed, int conjugated, int on_device) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(x, "x");
        check_ptr(y, "y");
        OpHost &o = op->o;
        if (ncols < 0) fail(QS_EDIM, "ncols must be non-negative");
        if (o.is_complex && !dtype_complex(dtype))
            fail(QS_EARG, "a complex operator needs a complex vector dtype");
        hipStream_t s = ensure_stream(o);
        const int64_t n = o.basis->Ns;
        const size_t esz = dtype_size(dtype), bytes = (size_t)n * ncols * esz;
        const void *dx = x;
        void *dy = y;
        bool beta_zero = beta_re == 0 && beta_im == 0;
        if (!on_device) {
            ensure_stage(o.stage_x, bytes);
            ensure_stage(o.stage_y, bytes);
            QS_HIP(hipMemcpyAsync(o.stage_x.get(), x, bytes, hipMemcpyHostToDevice, s));
            if (!beta_zero) QS_HIP(hipMemcpyAsync(o.stage_y.get(), y, bytes, hipMemcpyHostToDevice, s));
            dx = o.stage_x.get();
            dy = o.stage_y.get();
        } else if (x == y) {
            // the reference copies an aliased input (operators.jl:3158)
            ensure_stage(o.stage_x, bytes);
            QS_HIP(hipMemcpyAsync(o.stage_x.get(), x, bytes, hipMemcpyDeviceToDevice, s));
            dx = o.stage_x.get();
        }
        // M x, M^dagger x, M^T x = conj(M^dagger conj x), conj(M) x = conj(M conj x)
        bool adj = transposed != 0;
        bool cj = (transposed != 0) != (conjugated != 0);
        double ar = alpha_re, ai = cj ? -alpha_im : alpha_im, br = beta_re, bi = cj ? -beta_im : beta_im;
        // without symmetry a conjugated product conjugates alpha with the
        // coupling amplitudes but not on the extracted diagonal
        // (operators.jl:2581 vs :3097-3117): apply everything with conj(alpha),
        // then add (alpha - conj(alpha)) times the extracted diagonal
        const bool split = conjugated && alpha_im != 0 && o.extracted;
        for (int64_t c = 0; c < ncols; c++) {
            const char *xc = (const char *)dx + (size_t)c * n * esz;
            char *yc = (char *)dy + (size_t)c * n * esz;
            if (!split) {
                launch_matvec(o, dtype, xc, yc, ar, ai, br, bi, adj, cj, cj, s);
            } else {
                launch_matvec(o, dtype, xc, yc, ar, -ai, br, bi, adj, cj, cj, s);
                launch_matvec(*o.extracted, dtype, xc, yc, 0.0, 2 * ai, 1.0, 0.0, adj, cj, cj, s);
            }
        }
        if (!on_device) {
            QS_HIP(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, s));
            QS_HIP(hipStreamSynchronize(s));
        }
    });
}

int qs_op_matvec_noalloc(qs_op *op, int dtype, const void *x, void *y, int64_t ncols, double alpha_re,
                         double alpha_im, double beta_re, double beta_im, int transposed, int conjugated) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(x, "x");
        check_ptr(y, "y");
        OpHost &o = op->o;
        if (ncols < 0) fail(QS_EDIM, "ncols must be non-negative");
        if (dtype != QS_F64 && dtype != QS_C128 && dtype != QS_F32 && dtype != QS_C64) fail(QS_EARG, "unknown dtype");
        if (o.is_complex && !dtype_complex(dtype))
            fail(QS_EARG, "a complex operator needs a complex vector dtype");
        if (x == y)
            fail(QS_EARG, "x and y must not alias: the copy of an aliased input needs a buffer (use qs_op_matvec)");
        const BasisHost &b = *o.basis;
        const int64_t n = b.Ns;
        // the parent-table pre-scale pass writes one FP64 (complex) value per row
        const size_t need = b.ptab.get() ? (size_t)n * (dtype_complex(dtype) ? 16 : 8) : 0;
        if (need > QS_NOALLOC_SCRATCH_BYTES)
            fail(QS_EARG, "sector of " + std::to_string(n) + " rows needs " + std::to_string(need) +
                              " bytes of scratch, more than the no-allocation maximum of " +
                              std::to_string((size_t)QS_NOALLOC_SCRATCH_BYTES) + " bytes (use qs_op_matvec)");
        hipStream_t s = o.stream ? o.stream : o.own_stream;
        if (!s) fail(QS_EDEVICE, "operator has no stream");
        void *scr = noalloc_scratch();
        const size_t esz = dtype_size(dtype);
        bool adj = transposed != 0;
        bool cj = (transposed != 0) != (conjugated != 0);
        double ar = alpha_re, ai = cj ? -alpha_im : alpha_im, br = beta_re, bi = cj ? -beta_im : beta_im;
        const bool split = conjugated && alpha_im != 0 && o.extracted;
        for (int64_t c = 0; c < ncols; c++) {
            const char *xc = (const char *)x + (size_t)c * n * esz;
            char *yc = (char *)y + (size_t)c * n * esz;
            if (!split) {
                launch_matvec(o, dtype, xc, yc, ar, ai, br, bi, adj, cj, cj, s, nullptr, scr);
            } else {
                launch_matvec(o, dtype, xc, yc, ar, -ai, br, bi, adj, cj, cj, s, nullptr, scr);
                launch_matvec(*o.extracted, dtype, xc, yc, 0.0, 2 * ai, 1.0, 0.0, adj, cj, cj, s, nullptr, scr);
            }
        }
    });
}

int qs_eigsh(qs_op *op, int dtype, int k, int which, double tol, int maxiter, int ncv, const void *v0, double *vals,
             void *vecs, int *nconv, int *n_matvec) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(vals, "vals");
        eigsh_host(op->o, ensure_stream(op->o), dtype, k, which, tol, maxiter, ncv, v0, vals, vecs, nconv, n_matvec);
    });
}
int qs_expm_multiply(qs_op *op, int dtype, const void *v, int64_t n_times, const double *scales, double tol,
                     int krylov_dim, void *out, int *n_matvec) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(v, "v");
        check_ptr(out, "out");
        expm_host(op->o, ensure_stream(op->o), dtype, v, n_times, scales, tol, krylov_dim, out, n_matvec);
    });
}

int qs_evolve_dynamic(qs_op *h0, int n_dyn, qs_op *const *dyn, qs_drive_fn drive, void *user, int dtype,
                      const void *v0, int64_t ncols, double t0, int64_t n_times, const double *times, int imag_time,
                      double rtol, double atol, double max_step, void *out, int64_t *n_steps, int64_t *n_matvec) {
    return guard([&] {
        check_ptr(h0, "h0");
        check_ptr(v0, "v0");
        if (n_dyn < 0) fail(QS_EARG, "n_dyn must be non-negative");
        if (n_times < 0) fail(QS_EARG, "n_times must be non-negative");
        if (n_times > 0) {
            check_ptr(times, "times");
            check_ptr(out, "out");
        }
        std::vector<const OpHost *> d;
        for (int j = 0; j < n_dyn; j++) {
            check_ptr(dyn[j], "dyn[j]");
            d.push_back(&dyn[j]->o);
        }
        ensure_stream(h0->o);
        std::vector<double> ts(times, times + n_times);
        evolve_dynamic_host(h0->o, d, drive, user, dtype, v0, ncols, t0, ts, imag_time, rtol, atol, max_step, out,
                            n_steps, n_matvec);
    });
}

int qs_lanczos(qs_op *op, int dtype, const void *v0, int m, int full_ortho, double eps, double *alphas,
               double *betas, double *E, double *V, void *Q_T, int *used) {
    return guard([&] {
        check_ptr(op, "op");
        check_ptr(v0, "v0");
        check_ptr(used, "used");
        OpHost &o = op->o;
        if (o.is_complex && !dtype_complex(dtype)) fail(QS_EARG, "a complex operator needs a complex vector dtype");
        hipStream_t s = ensure_stream(o);
        StreamScope scope(s);
        const int64_t n = o.basis->Ns;
        const size_t esz = dtype_size(dtype);
        if (!(1 <= m && m < n)) fail(QS_EARG, "Lanczos dimension m must satisfy 1 <= m < length(v0)");
        DevBuf<char> dv((size_t)n * esz), Q((size_t)n * m * esz);
        QS_HIP(hipMemcpyAsync(dv.get(), v0, (size_t)n * esz, hipMemcpyHostToDevice, s));
        std::vector<double> a, b;
        int u = 0;
        lanczos_device(o, dtype, dv.get(), m, full_ortho != 0, eps, Q.get(), a, b, u);
        *used = u;
        if (alphas) std::copy(a.begin(), a.end(), alphas);
        if (betas) std::copy(b.begin(), b.end(), betas);
        // eigen(SymTridiagonal(alphas, betas))
        std::vector<double> T((size_t)u * u, 0.0), w, Vv;
        for (int i = 0; i < u; i++) {
            T[(size_t)i * u + i] = a[i];
            if (i + 1 < u) T[(size_t)i * u + i + 1] = T[(size_t)(i + 1) * u + i] = b[i];
        }
        sym_eig(u, T, w, Vv);
        if (E) std::copy(w.begin(), w.end(), E);
        if (V)
            for (int i = 0; i < u; i++)
                for (int j = 0; j < u; j++) V[(size_t)j * u + i] = Vv[(size_t)i * u + j];
        if (Q_T) {
            QS_HIP(hipMemcpyAsync(Q_T, Q.get(), (size_t)n * u * esz, hipMemcpyDeviceToHost, s));
            QS_HIP(hipStreamSynchronize(s));
        }
    });
}

}  // extern "C"

struct qs_batch {
    std::vector<const qs::OpHost *> ops;
    int dtype = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<qs::BatchMatvec, void (*)(qs::BatchMatvec *)> bm{nullptr, [](qs::BatchMatvec *) {}};
    ~qs_batch() {
        bm.reset();   // its job table waits on the stream before freeing
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int qs_basis_parent_size(const qs_basis *b, int kind, int64_t *n) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(n, "n");
        if (kind != 0 && kind != 1) fail(QS_EARG, "parent kind must be 0 (own parent) or 1 (all states)");
        *n = parent_length(b->b, kind);
    });
}

int qs_basis_project(const qs_basis *b, int dtype, int kind, const void *parent, void *out) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(parent, "parent");
        check_ptr(out, "out");
        basis_project(b->b, dtype, kind, parent, out, nullptr);
        QS_HIP(hipStreamSynchronize(nullptr));
    });
}

int qs_basis_expand(const qs_basis *b, int dtype, int kind, const void *x, void *parent, int64_t ncols,
                    int accumulate) {
    return guard([&] {
        check_ptr(b, "basis");
        check_ptr(x, "x");
        check_ptr(parent, "parent");
        if (ncols < 0 || ncols > 0x7fffffff) fail(QS_EARG, "ncols out of range");
        basis_expand(b->b, dtype, kind, x, parent, (int)ncols, accumulate != 0, nullptr);
        QS_HIP(hipStreamSynchronize(nullptr));
    });
}

int qs_batch_create(qs_op *const *ops, int n_ops, int dtype, qs_batch **out) {
    return guard([&] {
        check_ptr(out, "out");
        *out = nullptr;
        if (n_ops <= 0) fail(QS_EARG, "a batch needs at least one operator");
        check_ptr(ops, "ops");
        auto h = std::make_unique<qs_batch>();
        for (int i = 0; i < n_ops; i++) {
            check_ptr(ops[i], "ops[i]");
            if (ops[i]->o.basis->wide) fail(QS_EARG, "the batch engine does not take wide bases");
            h->ops.push_back(&ops[i]->o);
        }
        dtype_size(dtype);
        h->dtype = dtype;
        h->bm = make_batch_matvec(h->ops, dtype);
        QS_HIP(hipStreamCreate(&h->stream));
        *out = h.release();
    });
}

int qs_blocks_evolve(qs_op *const *ops, const int32_t *kind, int n_ops, int dtype, const void *psi, int64_t n_times,
                     const double *scales, double tol, int krylov_dim, void *out, int32_t *n_matvec) {
    return guard([&] {
        if (n_ops <= 0) fail(QS_EARG, "BlockOps needs at least one block");
        check_ptr(ops, "ops");
        check_ptr(kind, "kind");
        check_ptr(psi, "psi");
        check_ptr(out, "out");
        if (!dtype_complex(dtype)) fail(QS_EARG, "block evolution needs a complex vector dtype");
        if (n_times < 0 || (n_times > 0 && !scales)) fail(QS_EARG, "bad time grid");
        const size_t esz = dtype_size(dtype);
        int64_t npar = -1;
        for (int j = 0; j < n_ops; j++) {
            check_ptr(ops[j], "ops[j]");
            const int64_t p = parent_length(*ops[j]->o.basis, kind[j]);
            if (npar >= 0 && p != npar) fail(QS_EDIM, "blocks of different parent spaces");
            npar = p;
        }
        hipStream_t s;
        QS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};
        DevBuf<char> dpsi((size_t)npar * esz), dout((size_t)npar * std::max<int64_t>(1, n_times) * esz);
        QS_HIP(hipMemcpyAsync(dpsi.get(), psi, (size_t)npar * esz, hipMemcpyHostToDevice, s));
        QS_HIP(hipMemsetAsync(dout.get(), 0, (size_t)npar * std::max<int64_t>(1, n_times) * esz, s));
        // _active_block_entries: project, keep blocks with |P' psi| > 1000 eps
        std::vector<DevBuf<char>> x(n_ops), st;
        std::vector<const OpHost *> act_ops;
        std::vector<int> act;
        VecWork wk(s);
        for (int j = 0; j < n_ops; j++) {
            const BasisHost &b = *ops[j]->o.basis;
            x[j].alloc((size_t)b.Ns * esz);
            basis_project(b, dtype, kind[j], dpsi.get(), x[j].get(), s);
            if (std::sqrt(vec_norm2(dtype, x[j].get(), b.Ns, wk, s)) > 1000 * 2.220446049250313e-16) {
                act.push_back(j);
                act_ops.push_back(&ops[j]->o);
            }
        }
        if (act.empty()) fail(QS_EARG, "initial state has no projection onto the selected blocks");
        std::vector<cd> sc((size_t)n_times);
        for (int64_t t = 0; t < n_times; t++) sc[t] = cd(scales[2 * t], scales[2 * t + 1]);
        std::vector<ExpmTask> tasks;
        for (size_t i = 0; i < act.size(); i++) {
            const int64_t ns = act_ops[i]->basis->Ns;
            st.emplace_back((size_t)ns * std::max<int64_t>(1, n_times) * esz);
            tasks.push_back(ExpmTask{(int)i, x[act[i]].get(), sc, st.back().get()});
        }
        std::vector<int> nmv;
        expm_times_batch(act_ops, dtype, tasks, tol, krylov_dim, s, nmv);
        for (size_t i = 0; i < act.size(); i++)
            basis_expand(*act_ops[i]->basis, dtype, kind[act[i]], st[i].get(), dout.get(), (int)n_times, true, s);
        QS_HIP(hipMemcpyAsync(out, dout.get(), (size_t)npar * n_times * esz, hipMemcpyDeviceToHost, s));
        QS_HIP(hipStreamSynchronize(s));
        if (n_matvec) {
            for (int j = 0; j < n_ops; j++) n_matvec[j] = 0;
            for (size_t i = 0; i < act.size(); i++) n_matvec[act[i]] = nmv[i];
        }
    });
}

int qs_batch_destroy(qs_batch *b) {
    return guard([&] { delete b; });
}

int qs_batch_matvec(qs_batch *b, int n_jobs, const int32_t *op_index, const void *const *x, void *const *y) {
    return guard([&] {
        check_ptr(b, "batch");
        if (n_jobs < 0) fail(QS_EARG, "n_jobs must be non-negative");
        std::vector<MatvecJob> jobs;
        for (int j = 0; j < n_jobs; j++) {
            if (op_index[j] < 0 || op_index[j] >= (int)b->ops.size()) fail(QS_EARG, "op_index out of range");
            check_ptr(x[j], "x[j]");
            check_ptr(y[j], "y[j]");
            jobs.push_back({op_index[j], x[j], y[j]});
        }
        if (!jobs.empty()) launch_batch_matvec(*b->bm, jobs, b->stream);
        QS_HIP(hipStreamSynchronize(b->stream));
    });
}

int qs_batch_expm_multiply(qs_batch *b, int n_jobs, const int32_t *op_index, const void *const *v,
                           const int64_t *n_times, const double *scales, double tol, int krylov_dim,
                           void *const *out, int32_t *n_matvec) {
    return guard([&] {
        check_ptr(b, "batch");
        if (n_jobs < 0) fail(QS_EARG, "n_jobs must be non-negative");
        std::vector<ExpmTask> tasks;
        size_t k = 0;
        for (int j = 0; j < n_jobs; j++) {
            if (op_index[j] < 0 || op_index[j] >= (int)b->ops.size()) fail(QS_EARG, "op_index out of range");
            if (n_times[j] < 0) fail(QS_EARG, "n_times must be non-negative");
            check_ptr(v[j], "v[j]");
            check_ptr(out[j], "out[j]");
            ExpmTask t;
            t.op = op_index[j];
            t.v = v[j];
            t.out = out[j];
            for (int64_t i = 0; i < n_times[j]; i++, k++) t.scales.push_back(cd(scales[2 * k], scales[2 * k + 1]));
            tasks.push_back(std::move(t));
        }
        std::vector<int> nmv;
        expm_times_batch(b->ops, b->dtype, tasks, tol, krylov_dim, b->stream, nmv);
        if (n_matvec)
            for (int j = 0; j < n_jobs; j++) n_matvec[j] = nmv[j];
    });
}

}  // extern "C"
