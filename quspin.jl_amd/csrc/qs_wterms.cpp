// qs_wterms.cpp -- operator term lists on bases of up to 256 sites.
//
// The couplings compile to the same monomial form as on 64-bit states
// (operators.jl:214-263 consolidation, :343-379 local operators; basis.jl:
// 1231-1259 for the wide states): a term is alive when (s & cmask) == cval,
// maps s to s ^ flip and has amplitude A (-1)^popcount(s & smask).  The masks
// are four words.  Term lists the group does not leave invariant act through
// their group average (P'HP = P' H_avg P).
#include <algorithm>
#include <cstring>
#include <map>
#include <set>

#include "qs_internal.h"

namespace qs {
namespace {

WState wpermute(const std::vector<int> &perm, const WState &u) {
    WState r;
    for (size_t i = 0; i < perm.size(); i++)
        if (u.bit((int)i)) r.set(perm[i]);
    return r;
}

struct WKey {
    WState flip, cmask, cval, smask;
    bool operator<(const WKey &o) const {
        if (flip != o.flip) return flip < o.flip;
        if (cmask != o.cmask) return cmask < o.cmask;
        if (cval != o.cval) return cval < o.cval;
        return smask < o.smask;
    }
};
WKey key_of(const WTerm &t) { return WKey{t.flip, t.cmask, t.cval, t.smask}; }

std::vector<WTerm> compile_wide(const std::vector<Coupling> &cs, int L, int pauli) {
    std::vector<WTerm> out;
    for (const Coupling &c : cs) {
        if (c.op.empty()) fail(QS_EARG, "operator string cannot be empty");
        if (c.sites.size() != c.op.size()) fail(QS_EARG, "operator arity and coupling sites differ");
        for (int s : c.sites)
            if (s < 1 || s > L) fail(QS_EARG, "site " + std::to_string(s) + " lies outside 1:" + std::to_string(L));
        {
            bool al;
            int o;
            cd a;
            for (char ch : c.op) wide_local(ch, 0, pauli, al, o, a);   // validates the characters
        }
        WTerm t;
        t.A = c.coef;
        bool dead = false;
        std::set<int> seen;
        for (size_t i = 0; i < c.sites.size() && !dead; i++) {
            const int site = c.sites[i];
            if (seen.count(site)) continue;
            seen.insert(site);
            // the operators on this site, right to left, on each input bit
            bool alive[2];
            int outb[2];
            cd amp[2];
            for (int b0 = 0; b0 < 2; b0++) {
                alive[b0] = true;
                outb[b0] = b0;
                amp[b0] = cd(1, 0);
                for (int j = (int)c.op.size() - 1; j >= 0 && alive[b0]; j--) {
                    if (c.sites[(size_t)j] != site) continue;
                    bool al;
                    int o;
                    cd a;
                    wide_local(c.op[(size_t)j], outb[b0], pauli, al, o, a);
                    if (!al) {
                        alive[b0] = false;
                        break;
                    }
                    outb[b0] = o;
                    amp[b0] *= a;
                }
            }
            const int bit = site - 1;
            if (!alive[0] && !alive[1]) {
                dead = true;
                break;
            }
            if (alive[0] && alive[1]) {
                const bool flip = outb[0] == 1;
                if (flip != (outb[1] == 0)) fail(QS_EARG, "non-monomial local operator");
                if (flip) t.flip.set(bit);
                const cd r = amp[1] / amp[0];
                if (std::abs(r - cd(1, 0)) < 1e-15) {
                } else if (std::abs(r + cd(1, 0)) < 1e-15) {
                    t.smask.set(bit);
                } else {
                    fail(QS_EARG, "unsupported local operator product");
                }
                t.A *= amp[0];
            } else {
                const int v = alive[0] ? 0 : 1;
                t.cmask.set(bit);
                if (v) t.cval.set(bit);
                if (outb[v] != v) t.flip.set(bit);
                t.A *= amp[v];
            }
        }
        if (dead || t.A == cd(0, 0)) continue;
        out.push_back(t);
    }
    return out;
}

// first-seen order, amplitudes of equal monomials summed, zeros dropped
std::vector<WTerm> consolidate_wide(const std::vector<WTerm> &t) {
    std::map<WKey, size_t> idx;
    std::vector<WTerm> out;
    for (const WTerm &x : t) {
        auto it = idx.find(key_of(x));
        if (it == idx.end()) {
            idx[key_of(x)] = out.size();
            out.push_back(x);
        } else {
            out[it->second].A += x.A;
        }
    }
    std::vector<WTerm> kept;
    for (auto &x : out)
        if (x.A != cd(0, 0)) kept.push_back(x);
    return kept;
}

// g O g^-1 for the group element e: u -> perm(u) ^ c
WTerm image(const WTerm &x, const GroupHost &g, size_t e) {
    const std::vector<int> &pm = g.eperm[e];
    const WState &c = g.ecw[e];
    WTerm o;
    o.flip = wpermute(pm, x.flip);
    o.cmask = wpermute(pm, x.cmask);
    o.cval = wpermute(pm, x.cval) ^ (c & o.cmask);
    o.smask = wpermute(pm, x.smask);
    o.A = x.A;
    if ((c & o.smask).popcount() & 1) o.A = -o.A;
    return o;
}

bool same_terms(const std::vector<WTerm> &a, const std::vector<WTerm> &b, double tol) {
    if (a.size() != b.size()) return false;
    std::map<WKey, cd> m;
    for (auto &x : a) m[key_of(x)] = x.A;
    for (auto &x : b) {
        auto it = m.find(key_of(x));
        if (it == m.end() || std::abs(it->second - x.A) > tol * std::max(1.0, std::abs(x.A))) return false;
    }
    return true;
}

std::vector<WTerm> symmetrize_wide(const std::vector<WTerm> &base, const GroupHost &g) {
    const size_t n = g.eperm.size();
    bool inv = true;
    for (size_t e = 0; e < n && inv; e++) {
        std::vector<WTerm> img;
        for (auto &x : base) img.push_back(image(x, g, e));
        inv = same_terms(consolidate_wide(img), base, 0.0);
    }
    if (inv) return base;
    std::vector<WTerm> all;
    const double w = 1.0 / (double)n;
    for (size_t e = 0; e < n; e++)
        for (auto &x : base) {
            WTerm y = image(x, g, e);
            y.A *= w;
            all.push_back(y);
        }
    return consolidate_wide(all);
}

// adjoint of a compiled coupling (same flip, shifted condition, conj amplitude)
WTerm adjoint_wide(const WTerm &t) {
    WTerm a = t;
    a.cval = t.cval ^ (t.flip & t.cmask);
    a.A = std::conj(t.A);
    if ((t.flip & t.smask).popcount() & 1) a.A = -a.A;
    return a;
}

void split_wide(const std::vector<WTerm> &all, std::vector<WTerm> &diag, std::vector<WTerm> &off,
                std::vector<WGroupDev> &groups) {
    diag.clear();
    off.clear();
    groups.clear();
    std::vector<WState> order;
    std::map<WState, std::vector<WTerm>> by;
    for (auto &x : all) {
        if (x.flip.zero()) {
            diag.push_back(x);
            continue;
        }
        if (!by.count(x.flip)) order.push_back(x.flip);
        by[x.flip].push_back(x);
    }
    for (const WState &f : order) {
        WGroupDev gr;
        for (int k = 0; k < 4; k++) gr.flip[k] = f.w[k];
        gr.t0 = (int)off.size();
        for (auto &x : by[f]) off.push_back(x);
        gr.t1 = (int)off.size();
        groups.push_back(gr);
    }
}

void upload_wterms(const std::vector<WTerm> &src, DevBuf<WTermDev> &dst) {
    std::vector<WTermDev> h(std::max<size_t>(1, src.size()));
    memset(h.data(), 0, sizeof(WTermDev) * h.size());
    for (size_t i = 0; i < src.size(); i++) {
        for (int k = 0; k < 4; k++) {
            h[i].cmask[k] = src[i].cmask.w[k];
            h[i].cval[k] = src[i].cval.w[k];
            h[i].smask[k] = src[i].smask.w[k];
        }
        h[i].are = src[i].A.real();
        h[i].aim = src[i].A.imag();
    }
    dst.alloc(h.size());
    QS_HIP(hipMemcpy(dst.get(), h.data(), sizeof(WTermDev) * h.size(), hipMemcpyHostToDevice));
}

void upload_wgroups(const std::vector<WGroupDev> &src, DevBuf<WGroupDev> &dst) {
    std::vector<WGroupDev> h(src);
    if (h.empty()) h.push_back(WGroupDev{{0, 0, 0, 0}, 0, 0});
    dst.alloc(h.size());
    QS_HIP(hipMemcpy(dst.get(), h.data(), sizeof(WGroupDev) * h.size(), hipMemcpyHostToDevice));
}

}  // namespace

void setup_wide_terms(OpHost &o, const std::vector<Coupling> &cs, bool check_symm) {
    const BasisHost &b = *o.basis;
    const std::vector<WTerm> H = consolidate_wide(compile_wide(cs, b.L, b.pauli));
    std::vector<WTerm> Havg = H;
    if (b.group.order() > 1) {
        Havg = symmetrize_wide(H, b.group);
        if (check_symm && !same_terms(Havg, H, 1e-12))
            fail(QS_EARG, "operator list violates the selected symmetry sector");
    }
    for (auto &x : Havg)
        if (x.A.imag() != 0) o.is_complex = 1;
    // pull-form kernel: plain products apply the adjoint couplings
    std::vector<WTerm> Hadj;
    for (auto &x : Havg) Hadj.push_back(adjoint_wide(x));
    split_wide(Hadj, o.wdiag, o.woff, o.wgroups);
    split_wide(Havg, o.wdiag_adj, o.woff_adj, o.wgroups_adj);
    upload_wterms(o.wdiag, o.d_wdiag);
    upload_wterms(o.woff, o.d_woff);
    upload_wgroups(o.wgroups, o.d_wgroups);
    upload_wterms(o.wdiag_adj, o.d_wdiag_adj);
    upload_wterms(o.woff_adj, o.d_woff_adj);
    upload_wgroups(o.wgroups_adj, o.d_wgroups_adj);
}

}  // namespace qs
