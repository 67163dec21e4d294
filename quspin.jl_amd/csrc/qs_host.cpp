// qs_host.cpp -- host-side logic of libqsgpu: symmetry groups, operator-term
// compilation, group averaging (the P'HP semantics), and small dense FP64
// linear algebra for the Krylov solvers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

#include <atomic>
#include <condition_variable>
#include <exception>
#include <mutex>
#include <thread>

#include <unistd.h>

#include "qs_internal.h"

namespace qs {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }
[[noreturn]] void fail(int code, const std::string &msg) { throw Error(code, msg); }
const std::string &last_error() { return g_last_error; }

uint64_t full_mask(int L) { return L >= 64 ? ~0ULL : ((1ULL << L) - 1); }

uint64_t bit_reverse(uint64_t u, int L)
{
    uint64_t r = 0;
    for (int i = 0; i < L; i++)
        if (u >> i & 1) r |= 1ULL << (L - 1 - i);
    return r;
}

uint64_t rotl(uint64_t u, int k, int L)
{
    k %= L;
    if (k < 0) k += L;
    if (k == 0) return u;
    return ((u << k) | (u >> (L - k))) & full_mask(L);
}

uint64_t apply_point(const GroupHost &g, int p, uint64_t u)
{
    uint64_t v = g.prev[p] ? bit_reverse(u, g.L) : u;
    return v ^ g.pflip[p];
}

int64_t binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    int64_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

GroupDev group_to_dev(const GroupHost &h)
{
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
        // T^t = rotl by t*a = rotr by o; the 32-bit window at offset s holds
        // that rotation in its top L bits when s + 32 - L = o (mod L)
        if (h.L <= 32) {
            int o = ((h.L - t * h.a) % h.L + h.L) % h.L;
            d.win[t] = ((o + h.L - 32) % h.L + h.L) % h.L;
        }
    }
    return d;
}

// --------------------------------------------------------------- groups --
namespace {

// A group element as an affine map on bits: u -> perm(u) ^ c.
struct Affine {
    std::vector<int> perm;   // bit i -> bit perm[i]
    uint64_t c = 0;
    bool operator<(const Affine &o) const { return std::tie(perm, c) < std::tie(o.perm, o.c); }
    bool operator==(const Affine &o) const { return perm == o.perm && c == o.c; }
};

uint64_t permute_bits(const std::vector<int> &perm, uint64_t u)
{
    uint64_t r = 0;
    for (size_t i = 0; i < perm.size(); i++)
        if (u >> i & 1) r |= 1ULL << perm[i];
    return r;
}

Affine compose(const Affine &g1, const Affine &g2)   // g1 o g2
{
    Affine r;
    r.perm.resize(g1.perm.size());
    for (size_t i = 0; i < g1.perm.size(); i++) r.perm[i] = g1.perm[g2.perm[i]];
    r.c = permute_bits(g1.perm, g2.c) ^ g1.c;
    return r;
}

Affine identity_map(int L)
{
    Affine a;
    a.perm.resize(L);
    for (int i = 0; i < L; i++) a.perm[i] = i;
    return a;
}

bool close_to(cd a, cd b) { return std::abs(a - b) < 1e-9; }

cd exact_root(int k, int n)
{
    // exp(2 pi i k / n) with exact values at the real and imaginary axes
    k %= n;
    if (k < 0) k += n;
    if (k == 0) return cd(1, 0);
    if (2 * k == n) return cd(-1, 0);
    if (4 * k == n) return cd(0, 1);
    if (4 * k == 3 * n) return cd(0, -1);
    double ang = 2.0 * M_PI * (double)k / (double)n;
    return cd(std::cos(ang), std::sin(ang));
}

}  // namespace

void build_group(BasisHost &b)
{
    const int L = b.L;
    const int kb = b.blocks[0];
    const char *names[6] = {"kblock", "pblock", "zblock", "pzblock", "zAblock", "zBblock"};
    bool requested = false;
    for (int i = 0; i < 6; i++) requested |= b.blocks[i] != QS_UNSET;
    GroupHost &g = b.group;
    g = GroupHost();
    g.L = L;
    g.a = b.a;
    g.lam[0] = cd(1, 0);
    g.pchar[0] = cd(1, 0);
    if (!requested) {
        b.mode = MODE_FULL;
        return;
    }
    // basis.jl:686-687, 717-723, 744-745
    if (!(b.a > 0 && L % b.a == 0)) fail(QS_EARG, "a must be a positive divisor of L");
    const int order = L / b.a;
    if (b.blocks[1] != QS_UNSET && kb != QS_UNSET) {
        int m = ((kb % order) + order) % order;
        if (!(m == 0 || (order % 2 == 0 && m == order / 2)))
            fail(QS_EARG, "pblock can be combined with kblock only at k=0 or pi");
    }
    for (int i = 1; i < 6; i++)
        if (b.blocks[i] != QS_UNSET && b.blocks[i] != 1 && b.blocks[i] != -1)
            fail(QS_EARG, std::string(names[i]) + " must be +1 or -1");
    // generators that must keep the particle sector (symmetry_basis.jl:48-50,79-81)
    uint64_t odd = 0, even = 0;
    for (int site = 1; site <= L; site += 2) odd |= 1ULL << (site - 1);
    for (int site = 2; site <= L; site += 2) even |= 1ULL << (site - 1);
    if (b.nup >= 0) {
        bool z_ok = 2 * b.nup == L;
        if ((b.blocks[2] != QS_UNSET || b.blocks[3] != QS_UNSET) && !z_ok)
            fail(QS_EARG, "symmetry transformation leaves the selected particle sector");
        if (b.blocks[4] != QS_UNSET || b.blocks[5] != QS_UNSET)
            fail(QS_EARG, "symmetry transformation leaves the selected particle sector");
    }
    if ((b.blocks[4] != QS_UNSET || b.blocks[5] != QS_UNSET) && kb != QS_UNSET && (b.a % 2 == 1))
        fail(QS_EARG, "zAblock/zBblock combined with kblock need an even translation step a");

    // generators with their characters
    std::vector<std::pair<Affine, cd>> gens;
    Affine rev = identity_map(L);
    for (int i = 0; i < L; i++) rev.perm[i] = L - 1 - i;
    if (kb != QS_UNSET) {
        Affine t = identity_map(L);
        for (int i = 0; i < L; i++) t.perm[i] = (i + b.a) % L;
        gens.push_back({t, exact_root(kb, order)});
    }
    if (b.blocks[1] != QS_UNSET) gens.push_back({rev, cd(b.blocks[1], 0)});
    if (b.blocks[2] != QS_UNSET) {
        Affine z = identity_map(L);
        z.c = full_mask(L);
        gens.push_back({z, cd(b.blocks[2], 0)});
    }
    if (b.blocks[3] != QS_UNSET) {
        Affine pz = rev;
        pz.c = full_mask(L);
        gens.push_back({pz, cd(b.blocks[3], 0)});
    }
    if (b.blocks[4] != QS_UNSET) {
        Affine za = identity_map(L);
        za.c = odd;
        gens.push_back({za, cd(b.blocks[4], 0)});
    }
    if (b.blocks[5] != QS_UNSET) {
        Affine zb = identity_map(L);
        zb.c = even;
        gens.push_back({zb, cd(b.blocks[5], 0)});
    }
    // closure by breadth-first search with character bookkeeping
    std::map<Affine, cd> chi;
    std::vector<Affine> queue{identity_map(L)};
    chi[queue[0]] = cd(1, 0);
    bool consistent = true;
    for (size_t qi = 0; qi < queue.size(); qi++) {
        Affine cur = queue[qi];
        cd cc = chi[cur];
        for (auto &gg : gens) {
            Affine nx = compose(gg.first, cur);
            cd nc = gg.second * cc;
            auto it = chi.find(nx);
            if (it == chi.end()) {
                chi[nx] = nc;
                queue.push_back(nx);
                if (queue.size() > 4096) fail(QS_EARG, "symmetry group too large");
            } else if (!close_to(it->second, nc)) {
                consistent = false;
            }
        }
    }
    if (!consistent) fail(QS_EARG, "the requested symmetry sector is empty");

    // decompose every element as T^t o X_p
    g.nT = (kb != QS_UNSET) ? order : 1;
    for (int t = 0; t < g.nT; t++) g.lam[t] = (kb != QS_UNSET) ? exact_root((int64_t)kb * t % order, order) : cd(1, 0);
    std::vector<std::tuple<int, uint64_t, cd>> points;
    for (auto &kv : chi) {
        const Affine &e = kv.first;
        bool placed = false;
        for (int t = 0; t < g.nT && !placed; t++) {
            int sh = t * b.a;
            bool is_id = true, is_rev = true;
            for (int i = 0; i < L; i++) {
                int p = ((e.perm[i] - sh) % L + L) % L;
                is_id &= p == i;
                is_rev &= p == L - 1 - i;
            }
            if (!is_id && !is_rev) continue;
            placed = true;
            if (t != 0) continue;   // point ops are the t = 0 representatives
            int r = is_id ? 0 : 1;
            uint64_t f = e.c;
            points.emplace_back(r, f, kv.second);
        }
        if (!placed) fail(QS_EARG, "unsupported symmetry combination");
    }
    if ((int)points.size() > kMaxPoint) fail(QS_EARG, "symmetry group too large");
    // The kernels run over the pairs (t, p); when two pairs are the same map
    // (tiny L) their characters must agree, otherwise the projection is zero.
    for (auto &pt : points)
        for (int t = 0; t < g.nT; t++) {
            Affine x = identity_map(L);
            if (std::get<0>(pt))
                for (int i = 0; i < L; i++) x.perm[i] = L - 1 - i;
            x.c = std::get<1>(pt);
            Affine tr = identity_map(L);
            for (int i = 0; i < L; i++) tr.perm[i] = (i + t * b.a) % L;
            Affine e = compose(tr, x);
            cd want = (kb != QS_UNSET ? exact_root((int64_t)kb * t % order, order) : cd(1, 0)) * std::get<2>(pt);
            auto it = chi.find(e);
            if (it == chi.end() || !close_to(it->second, want))
                fail(QS_EARG, "the requested symmetry sector is empty");
        }
    std::sort(points.begin(), points.end(), [](auto &x, auto &y) {
        return std::make_pair(std::get<0>(x), std::get<1>(x)) < std::make_pair(std::get<0>(y), std::get<1>(y));
    });
    g.npoint = (int)points.size();
    for (int p = 0; p < g.npoint; p++) {
        g.prev[p] = std::get<0>(points[p]);
        g.pflip[p] = std::get<1>(points[p]);
        g.pchar[p] = std::get<2>(points[p]);
    }
    b.is_complex = 0;
    for (auto &kv : chi)
        if (std::abs(kv.second.imag()) > 1e-12) b.is_complex = 1;
    bool zonly = b.blocks[0] == QS_UNSET && b.blocks[1] == QS_UNSET && b.blocks[2] != QS_UNSET &&
                 b.blocks[3] == QS_UNSET && b.blocks[4] == QS_UNSET && b.blocks[5] == QS_UNSET;
    if (zonly) {
        b.mode = MODE_ZONLY;
        b.zchar = b.blocks[2];
    } else {
        b.mode = MODE_STORED;
    }
}

// basis.jl:530-620 (SpinBasisGeneral, spin 1/2, L <= 32 here): the group the
// blocks' maps generate, every element with its character (the product of
// the generators' eigenvalues along any word; a clash means no common
// eigenvector, so the sector is empty).
void build_group_general(BasisHost &b)
{
    const int L = b.L;
    GroupHost &g = b.group;
    g = GroupHost();
    g.L = L;
    g.a = 1;
    g.lam[0] = cd(1, 0);
    g.pchar[0] = cd(1, 0);
    if (b.gmaps.empty()) {
        b.mode = MODE_FULL;
        return;
    }
    if (L > 32)
        fail(QS_EARG, "general symmetry maps on more than 32 sites are not supported on the GPU");
    std::vector<std::pair<Affine, cd>> gens;
    for (const GenMap &m : b.gmaps) {
        uint64_t f = m.flip;
        // the particle sector is kept only by no flip, or by inverting every site at half filling
        if (b.nup >= 0 && f != 0 && !(f == full_mask(L) && 2 * b.nup == L))
            fail(QS_EARG, "symmetry transformation leaves the selected particle sector");
        Affine a;
        a.perm = m.perm;
        a.c = permute_bits(m.perm, f);
        gens.push_back({a, exact_root(((m.q % m.order) + m.order) % m.order, m.order)});
    }
    std::map<Affine, cd> chi;
    std::vector<Affine> queue{identity_map(L)};
    chi[queue[0]] = cd(1, 0);
    bool consistent = true;
    for (size_t qi = 0; qi < queue.size(); qi++) {
        const Affine cur = queue[qi];
        const cd cc = chi[cur];
        for (auto &gg : gens) {
            Affine nx = compose(gg.first, cur);
            cd nc = gg.second * cc;
            auto it = chi.find(nx);
            if (it == chi.end()) {
                chi[nx] = nc;
                queue.push_back(nx);
                if (queue.size() > (size_t)kMaxGeneral) fail(QS_EARG, "symmetry group too large");
            } else if (!close_to(it->second, nc)) {
                consistent = false;
            }
        }
    }
    if (!consistent) fail(QS_EARG, "the requested symmetry sector is empty");
    g.general = true;
    for (const Affine &e : queue) {
        g.eperm.push_back(e.perm);
        g.ec.push_back(e.c);
        g.echar.push_back(chi[e]);
    }
    b.is_complex = 0;
    for (const cd &c : g.echar)
        if (std::abs(c.imag()) > 1e-12) b.is_complex = 1;
    b.mode = MODE_STORED;
}

uint64_t group_apply(const GroupHost &g, int e, uint64_t u)
{
    if (g.general) return permute_bits(g.eperm[(size_t)e], u) ^ g.ec[(size_t)e];
    const int t = e / g.npoint, p = e - t * g.npoint;
    return rotl(apply_point(g, p, u), t * g.a, g.L);
}

cd group_char(const GroupHost &g, int e)
{
    if (g.general) return g.echar[(size_t)e];
    const int t = e / g.npoint, p = e - t * g.npoint;
    return g.lam[t] * g.pchar[p];
}

// ------------------------------------------------------------ wide groups --
namespace {

// A group element on up to 256 sites: u -> perm(u) ^ c.
struct WAffine {
    std::vector<int> perm;
    WState c;
    bool operator<(const WAffine &o) const {
        if (perm != o.perm) return perm < o.perm;
        return c < o.c;
    }
};

WState wpermute(const std::vector<int> &perm, const WState &u)
{
    WState r;
    for (size_t i = 0; i < perm.size(); i++)
        if (u.bit((int)i)) r.set(perm[i]);
    return r;
}

WAffine wcompose(const WAffine &g1, const WAffine &g2)   // g1 o g2
{
    WAffine r;
    r.perm.resize(g1.perm.size());
    for (size_t i = 0; i < g1.perm.size(); i++) r.perm[i] = g1.perm[g2.perm[i]];
    r.c = wpermute(g1.perm, g2.c) ^ g1.c;
    return r;
}

}  // namespace

// basis.jl:361-445 and 530-620 on more than 32 sites (up to 256): the group
// the blocks' maps generate, every element with its character, as in
// build_group_general but with four-word xor constants.
void build_group_wide(BasisHost &b)
{
    const int L = b.L;
    GroupHost &g = b.group;
    g = GroupHost();
    g.L = L;
    g.a = 1;
    g.lam[0] = cd(1, 0);
    g.pchar[0] = cd(1, 0);
    WState full;
    for (int i = 0; i < L; i++) full.set(i);
    std::vector<std::pair<WAffine, cd>> gens;
    for (const GenMap &m : b.gmaps) {
        const WState f = m.flipw;
        if (b.nup >= 0 && !f.zero() && !(f == full && 2 * b.nup == L))
            fail(QS_EARG, "symmetry transformation leaves the selected particle sector");
        WAffine a;
        a.perm = m.perm;
        a.c = wpermute(m.perm, f);
        gens.push_back({a, exact_root(((m.q % m.order) + m.order) % m.order, m.order)});
    }
    WAffine id;
    id.perm.resize(L);
    for (int i = 0; i < L; i++) id.perm[i] = i;
    std::map<WAffine, cd> chi;
    std::vector<WAffine> queue{id};
    chi[id] = cd(1, 0);
    bool consistent = true;
    for (size_t qi = 0; qi < queue.size(); qi++) {
        const WAffine cur = queue[qi];
        const cd cc = chi[cur];
        for (auto &gg : gens) {
            WAffine nx = wcompose(gg.first, cur);
            cd nc = gg.second * cc;
            auto it = chi.find(nx);
            if (it == chi.end()) {
                chi[nx] = nc;
                queue.push_back(nx);
                if (queue.size() > (size_t)kMaxGeneral) fail(QS_EARG, "symmetry group too large");
            } else if (!close_to(it->second, nc)) {
                consistent = false;
            }
        }
    }
    if (!consistent) fail(QS_EARG, "the requested symmetry sector is empty");
    g.general = true;
    for (const WAffine &e : queue) {
        g.eperm.push_back(e.perm);
        g.ecw.push_back(e.c);
        g.ec.push_back(e.c.w[0]);
        g.echar.push_back(chi[e]);
    }
    b.is_complex = 0;
    for (const cd &c : g.echar)
        if (std::abs(c.imag()) > 1e-12) b.is_complex = 1;
    b.mode = MODE_STORED;
}

WState wide_group_apply(const GroupHost &g, int e, const WState &u)
{
    return wpermute(g.eperm[(size_t)e], u) ^ g.ecw[(size_t)e];
}

// ----------------------------------------------------------------- terms --
namespace {

struct LocalAct {
    bool alive;
    int out;
    cd amp;
};

LocalAct local_op(char op, int b, int pauli)
{
    // operators.jl:343-379
    double sz = pauli ? 1.0 : 0.5, spm = pauli ? 2.0 : 1.0;
    switch (op) {
    case 'I': return {true, b, cd(1, 0)};
    case 'z': return {true, b, cd(b ? sz : -sz, 0)};
    case '+': return b ? LocalAct{false, b, cd(0, 0)} : LocalAct{true, 1, cd(spm, 0)};
    case '-': return b ? LocalAct{true, 0, cd(spm, 0)} : LocalAct{false, b, cd(0, 0)};
    case 'x': return {true, 1 - b, cd(sz, 0)};
    case 'y': return {true, 1 - b, b ? cd(0, -sz) : cd(0, sz)};
    }
    fail(QS_EARG, std::string("unsupported spin operator '") + op + "'");
}

}  // namespace

void wide_local(char op, int bit, int pauli, bool &alive, int &out, cd &amp)
{
    const LocalAct a = local_op(op, bit, pauli);
    alive = a.alive;
    out = a.out;
    amp = a.amp;
}

std::vector<CTerm> compile_couplings(const std::vector<Coupling> &cs, int L, int pauli)
{
    std::vector<CTerm> out;
    for (const Coupling &c : cs) {
        if (c.op.empty()) fail(QS_EARG, "operator string cannot be empty");
        if (c.sites.size() != c.op.size()) fail(QS_EARG, "operator arity and coupling sites differ");
        for (int s : c.sites)
            if (s < 1 || s > L)
                fail(QS_EARG, "site " + std::to_string(s) + " lies outside 1:" + std::to_string(L));
        for (char ch : c.op) local_op(ch, 0, pauli);   // validates the characters
        CTerm t{0, 0, 0, 0, c.coef};
        bool dead = false;
        std::set<int> seen;
        for (size_t i = 0; i < c.sites.size() && !dead; i++) {
            int site = c.sites[i];
            if (seen.count(site)) continue;
            seen.insert(site);
            LocalAct res[2];
            for (int b0 = 0; b0 < 2; b0++) {
                LocalAct st{true, b0, cd(1, 0)};
                for (int j = (int)c.op.size() - 1; j >= 0 && st.alive; j--) {
                    if (c.sites[j] != site) continue;
                    LocalAct la = local_op(c.op[j], st.out, pauli);
                    if (!la.alive) { st.alive = false; break; }
                    st.out = la.out;
                    st.amp *= la.amp;
                }
                res[b0] = st;
            }
            uint64_t bit = 1ULL << (site - 1);
            if (!res[0].alive && !res[1].alive) { dead = true; break; }
            if (res[0].alive && res[1].alive) {
                bool flip = res[0].out == 1;
                if (flip != (res[1].out == 0)) fail(QS_EARG, "non-monomial local operator");
                if (flip) t.flip |= bit;
                cd r = res[1].amp / res[0].amp;
                if (std::abs(r - cd(1, 0)) < 1e-15) {
                } else if (std::abs(r + cd(1, 0)) < 1e-15) {
                    t.smask |= bit;
                } else {
                    fail(QS_EARG, "unsupported local operator product");
                }
                t.A *= res[0].amp;
            } else {
                int v = res[0].alive ? 0 : 1;
                t.cmask |= bit;
                if (v) t.cval |= bit;
                if (res[v].out != v) t.flip |= bit;
                t.A *= res[v].amp;
            }
        }
        if (dead || t.A == cd(0, 0)) continue;
        out.push_back(t);
    }
    return out;
}

std::vector<CTerm> consolidate_cterms(const std::vector<CTerm> &t)
{
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>, size_t> idx;
    std::vector<CTerm> out;
    for (const CTerm &x : t) {
        auto key = std::make_tuple(x.flip, x.cmask, x.cval, x.smask);
        auto it = idx.find(key);
        if (it == idx.end()) {
            idx[key] = out.size();
            out.push_back(x);
        } else {
            out[it->second].A += x.A;
        }
    }
    std::vector<CTerm> kept;
    for (auto &x : out)
        if (x.A != cd(0, 0)) kept.push_back(x);
    return kept;
}

namespace {

// g O g^-1 for g(u) = pi(u) ^ f with pi = rotl(rev^r(.), sh)
CTerm conjugate_term(const CTerm &t, int r, uint64_t fx, int sh, int L)
{
    auto pi = [&](uint64_t v) { return rotl(r ? bit_reverse(v, L) : v, sh, L); };
    uint64_t f = rotl(fx, sh, L);
    CTerm o;
    o.flip = pi(t.flip);
    o.cmask = pi(t.cmask);
    o.cval = pi(t.cval) ^ (f & o.cmask);
    o.smask = pi(t.smask);
    o.A = t.A;
    if (__builtin_popcountll(f & o.smask) & 1) o.A = -o.A;
    return o;
}

bool same_dict(const std::vector<CTerm> &a, const std::vector<CTerm> &b)
{
    if (a.size() != b.size()) return false;
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>, cd> m;
    for (auto &x : a) m[std::make_tuple(x.flip, x.cmask, x.cval, x.smask)] = x.A;
    for (auto &x : b) {
        auto it = m.find(std::make_tuple(x.flip, x.cmask, x.cval, x.smask));
        if (it == m.end() || it->second != x.A) return false;
    }
    return true;
}

}  // namespace

std::vector<CTerm> symmetrize_cterms(const std::vector<CTerm> &t, const BasisHost &b)
{
    const GroupHost &g = b.group;
    if (b.mode == MODE_FULL) return consolidate_cterms(t);
    std::vector<CTerm> base = consolidate_cterms(t);
    if (g.general) {
        // the same test and average over the explicit elements
        auto image = [&](const CTerm &x, size_t e) {
            auto pi = [&](uint64_t v) { return permute_bits(g.eperm[e], v); };
            CTerm o;
            o.flip = pi(x.flip);
            o.cmask = pi(x.cmask);
            o.cval = pi(x.cval) ^ (g.ec[e] & o.cmask);
            o.smask = pi(x.smask);
            o.A = x.A;
            if (__builtin_popcountll(g.ec[e] & o.smask) & 1) o.A = -o.A;
            return o;
        };
        bool inv = true;
        for (size_t e = 0; e < g.eperm.size() && inv; e++) {
            std::vector<CTerm> img;
            for (auto &x : base) img.push_back(image(x, e));
            inv = same_dict(consolidate_cterms(img), base);
        }
        if (inv) return base;
        std::vector<CTerm> all;
        const double w = 1.0 / (double)g.eperm.size();
        for (size_t e = 0; e < g.eperm.size(); e++)
            for (auto &x : base) {
                CTerm y = image(x, e);
                y.A *= w;
                all.push_back(y);
            }
        return consolidate_cterms(all);
    }
    bool invariant = true;
    for (int p = 0; p < g.npoint && invariant; p++)
        for (int tt = 0; tt < g.nT && invariant; tt++) {
            std::vector<CTerm> img;
            for (auto &x : base) img.push_back(conjugate_term(x, g.prev[p], g.pflip[p], tt * g.a, g.L));
            invariant = same_dict(consolidate_cterms(img), base);
        }
    if (invariant) return base;
    // P'HP = P' H_avg P with H_avg the group average of H (the character
    // drops out), so the projected action of a non-invariant term list is the
    // action of its average.
    std::vector<CTerm> all;
    double w = 1.0 / (double)g.order();
    for (int p = 0; p < g.npoint; p++)
        for (int tt = 0; tt < g.nT; tt++)
            for (auto &x : base) {
                CTerm y = conjugate_term(x, g.prev[p], g.pflip[p], tt * g.a, g.L);
                y.A *= w;
                all.push_back(y);
            }
    return consolidate_cterms(all);
}

void split_terms(const std::vector<CTerm> &all, std::vector<CTerm> &diag, std::vector<CTerm> &off,
                 std::vector<GroupOfTerms> &groups)
{
    diag.clear();
    off.clear();
    groups.clear();
    std::vector<uint64_t> order;
    std::map<uint64_t, std::vector<CTerm>> by;
    for (auto &x : all) {
        if (x.flip == 0) {
            diag.push_back(x);
            continue;
        }
        if (!by.count(x.flip)) order.push_back(x.flip);
        by[x.flip].push_back(x);
    }
    for (uint64_t f : order) {
        GroupOfTerms gr;
        gr.flip = f;
        gr.lo = gr.hi = -1;
        if (__builtin_popcountll(f) == 2) {
            gr.lo = __builtin_ctzll(f);
            gr.hi = 63 - __builtin_clzll(f);
        }
        gr.t0 = (int)off.size();
        for (auto &x : by[f]) off.push_back(x);
        gr.t1 = (int)off.size();
        groups.push_back(gr);
    }
}

FastTerms compile_fast(const std::vector<CTerm> &diag, const std::vector<CTerm> &off,
                       const std::vector<GroupOfTerms> &groups, int L, int nup)
{
    FastTerms f;
    // two-bit hop tables: the summed amplitude of a group as a function of
    // its two bits, accumulated in term order as the generic kernel does
    f.hop_ok = L <= 32;
    for (const auto &g : groups) {
        if (g.lo < 0 || !f.hop_ok) { f.hop_ok = false; break; }
        HopDev h{};
        h.lo = g.lo;
        h.hi = g.hi;
        h.lowmask = (uint32_t)((1ULL << g.lo) - 1);
        h.between = (uint32_t)(((1ULL << g.hi) - 1) & ~((2ULL << g.lo) - 1));
        h.lo65 = g.lo * 65;
        h.hi65 = g.hi * 65;
        h.top = g.hi == L - 1;
        for (int k = g.t0; k < g.t1; k++)
            if ((off[k].cmask | off[k].smask) & ~g.flip) f.hop_ok = false;
        if (!f.hop_ok) break;
        for (int pat = 0; pat < 4; pat++) {
            const uint64_t s = ((uint64_t)(pat & 1) << g.lo) | ((uint64_t)(pat >> 1) << g.hi);
            double re = 0, im = 0;
            for (int k = g.t0; k < g.t1; k++) {
                const CTerm &t = off[k];
                if ((s & t.cmask) != t.cval) continue;
                const double sg = (__builtin_popcountll(s & t.smask) & 1) ? -1.0 : 1.0;
                re += sg * t.A.real();
                im += sg * t.A.imag();
            }
            h.re[pat] = re;
            h.im[pat] = im;
            // with a particle number, equal bits leave the sector
            const bool in_sector = nup < 0 || pat == 1 || pat == 2;
            if (in_sector && (re != 0.0 || im != 0.0)) h.nzmask |= 1 << pat;
        }
        h.sym = h.re[1] == h.re[2] && h.im[1] == h.im[2];
        f.hops.push_back(h);
    }
    if (!f.hop_ok) f.hops.clear();
    // sign terms on one site or on a pair at cyclic distance d, grouped by (d, A)
    for (const auto &t : diag) {
        const int nb = __builtin_popcountll(t.smask);
        if (t.cmask != 0 || nb < 1 || nb > 2) {
            f.drest.push_back(t);
            continue;
        }
        int d = 0, base = __builtin_ctzll(t.smask);
        if (nb == 2) {
            const int i = base, j = 63 - __builtin_clzll(t.smask);
            const int d1 = j - i;
            if (d1 <= L - d1) { d = d1; base = i; }
            else { d = L - d1; base = j; }
        }
        bool placed = false;
        for (auto &c : f.dcls) {
            if (c.d == d && c.re == t.A.real() && c.im == t.A.imag() && !((c.mask >> base) & 1)) {
                c.mask |= 1ULL << base;
                c.n++;
                placed = true;
                break;
            }
        }
        if (!placed) f.dcls.push_back(DiagClass{1ULL << base, d, 1, t.A.real(), t.A.imag()});
    }
    return f;
}

// ---------------------------------------------------- dense eigensolvers --
// Real symmetric n x n eigenproblem (row-major A, destroyed): Householder
// reduction to tridiagonal form, then implicit symmetric QR steps with the
// Wilkinson shift (Golub & Van Loan, Matrix Computations, algorithms 8.3.1
// and 8.3.3), the Givens rotations accumulated onto the reduction's
// orthogonal factor.  A matrix that is already tridiagonal (the Lanczos
// projections) skips the reduction.  Eigenvalues ascending in w,
// eigenvectors in the columns of V (row-major).
void sym_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V)
{
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    // the orthogonal factor is accumulated transposed (Zt row c = column c),
    // so each rotation and reflection streams contiguous rows
    std::vector<double> Zt((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Zt[(size_t)i * n + i] = 1.0;
    auto zrow = [&](int c) { return Zt.data() + (size_t)c * n; };
    V = Zt;
    w.assign(n, 0.0);
    if (n == 0) return;
    bool finite = true;
    for (double x : A) finite = finite && std::isfinite(x);
    if (!finite) {
        w.assign(n, std::nan(""));
        return;
    }
    // Householder reduction: A = Q T Q^T
    std::vector<double> v(n), p(n), t(n);
    for (int k = 0; k + 2 < n; k++) {
        const int m = n - k - 1;
        double sig = 0;
        for (int i = 0; i < m; i++) sig += a(k + 1 + i, k) * a(k + 1 + i, k);
        double tail = sig - a(k + 1, k) * a(k + 1, k);
        if (tail == 0) continue;   // column already reduced
        sig = std::sqrt(sig);
        const double x0 = a(k + 1, k);
        const double alpha = x0 >= 0 ? -sig : sig;
        for (int i = 0; i < m; i++) v[i] = a(k + 1 + i, k);
        v[0] = x0 - alpha;
        double vv = 0;
        for (int i = 0; i < m; i++) vv += v[i] * v[i];
        const double beta = 2.0 / vv;
        // trailing block: A22 -= v w^T + w v^T, p = beta A22 v, w = p - (beta p.v / 2) v
        double pv = 0;
        for (int i = 0; i < m; i++) {
            double acc = 0;
            for (int j = 0; j < m; j++) acc += a(k + 1 + i, k + 1 + j) * v[j];
            p[i] = beta * acc;
            pv += p[i] * v[i];
        }
        const double kf = 0.5 * beta * pv;
        for (int i = 0; i < m; i++) p[i] -= kf * v[i];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) a(k + 1 + i, k + 1 + j) -= v[i] * p[j] + p[i] * v[j];
        a(k + 1, k) = a(k, k + 1) = alpha;
        for (int i = 1; i < m; i++) a(k + 1 + i, k) = a(k, k + 1 + i) = 0.0;
        // Q <- Q (I - beta v v^T) on columns k+1 ..
        std::fill(t.begin(), t.end(), 0.0);
        for (int j = 0; j < m; j++) {
            const double *zr = zrow(k + 1 + j);
            for (int r = 0; r < n; r++) t[r] += v[j] * zr[r];
        }
        for (int j = 0; j < m; j++) {
            double *zr = zrow(k + 1 + j);
            const double f = beta * v[j];
            for (int r = 0; r < n; r++) zr[r] -= f * t[r];
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; i++) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; i++) e[i] = 0.5 * (a(i, i + 1) + a(i + 1, i));
    // implicit QR on the unreduced trailing block [lo, hi], bulge chased
    // down by rotations G = [[c, s], [-s, c]] on (k, k+1)
    const double eps = 2.220446049250313e-16;
    auto small = [&](int i) { return std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1])); };
    int hi = n - 1, sweeps = 0;
    while (hi > 0) {
        if (small(hi - 1)) {
            e[hi - 1] = 0;
            hi--;
            continue;
        }
        int lo = hi - 1;
        while (lo > 0 && !small(lo - 1)) lo--;
        if (lo > 0) e[lo - 1] = 0;
        if (++sweeps > 64 * n) break;   // not reached for finite input
        const double dl = 0.5 * (d[hi - 1] - d[hi]), eb = e[hi - 1];
        const double mu = d[hi] - eb * eb / (dl + std::copysign(std::hypot(dl, eb), dl));
        double x = d[lo] - mu, z = e[lo];
        for (int k = lo; k < hi; k++) {
            const double r = std::sqrt(x * x + z * z);
            const double c = r == 0 ? 1.0 : x / r, sn = r == 0 ? 0.0 : -z / r;
            if (k > lo) e[k - 1] = r;
            const double dk = d[k], dk1 = d[k + 1], ek = e[k];
            d[k] = c * c * dk - 2 * c * sn * ek + sn * sn * dk1;
            d[k + 1] = sn * sn * dk + 2 * c * sn * ek + c * c * dk1;
            e[k] = c * sn * (dk - dk1) + (c * c - sn * sn) * ek;
            if (k + 1 < hi) {
                z = -sn * e[k + 1];
                e[k + 1] *= c;
                x = e[k];
            }
            double *z0 = zrow(k), *z1 = zrow(k + 1);
            for (int r2 = 0; r2 < n; r2++) {
                const double zk = z0[r2], zk1 = z1[r2];
                z0[r2] = c * zk - sn * zk1;
                z1[r2] = sn * zk + c * zk1;
            }
        }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int j = 0; j < n; j++) {
        w[j] = d[ord[j]];
        const double *zr = zrow(ord[j]);
        for (int k = 0; k < n; k++) V[(size_t)k * n + j] = zr[k];
    }
}

// Complex Hermitian eigenproblem via the real symmetric embedding
// [Re -Im; Im Re] (each eigenvalue appears twice; one copy is kept).
void herm_eig(int n, std::vector<cd> &A, std::vector<double> &w, std::vector<cd> &V)
{
    int m = 2 * n;
    std::vector<double> B((size_t)m * m);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            cd z = A[(size_t)i * n + j];
            B[(size_t)i * m + j] = z.real();
            B[(size_t)i * m + j + n] = -z.imag();
            B[(size_t)(i + n) * m + j] = z.imag();
            B[(size_t)(i + n) * m + j + n] = z.real();
        }
    std::vector<double> ww, VV;
    sym_eig(m, B, ww, VV);
    // pick n orthonormal complex vectors: take pairs in order, orthogonalise
    w.clear();
    V.assign((size_t)n * n, cd(0, 0));
    int got = 0;
    for (int j = 0; j < m && got < n; j++) {
        std::vector<cd> v(n);
        for (int k = 0; k < n; k++) v[k] = cd(VV[(size_t)k * m + j], VV[(size_t)(k + n) * m + j]);
        for (int c = 0; c < got; c++) {
            cd d = 0;
            for (int k = 0; k < n; k++) d += std::conj(V[(size_t)k * n + c]) * v[k];
            for (int k = 0; k < n; k++) v[k] -= d * V[(size_t)k * n + c];
        }
        double nr = 0;
        for (int k = 0; k < n; k++) nr += std::norm(v[k]);
        nr = std::sqrt(nr);
        if (nr < 0.5) continue;
        for (int k = 0; k < n; k++) V[(size_t)k * n + got] = v[k] / nr;
        w.push_back(ww[j]);
        got++;
    }
}

// exp(A) for a small complex matrix: scaling and squaring with the [13/13]
// Pade approximant (Higham, SIAM J. Matrix Anal. Appl. 26 (2005) 1179).
void expm_dense(int n, const std::vector<cd> &A, std::vector<cd> &E)
{
    // complex products on split real / imaginary planes, which vectorise
    std::vector<double> yr((size_t)n * n), yi((size_t)n * n), zr((size_t)n * n), zi((size_t)n * n);
    auto mm = [&](const std::vector<cd> &X, const std::vector<cd> &Y) {
        for (size_t t = 0; t < (size_t)n * n; t++) {
            yr[t] = Y[t].real();
            yi[t] = Y[t].imag();
        }
        std::fill(zr.begin(), zr.end(), 0.0);
        std::fill(zi.begin(), zi.end(), 0.0);
        for (int i = 0; i < n; i++) {
            double *qr = zr.data() + (size_t)i * n, *qi = zi.data() + (size_t)i * n;
            for (int k = 0; k < n; k++) {
                const double a = X[(size_t)i * n + k].real(), b = X[(size_t)i * n + k].imag();
                if (a == 0 && b == 0) continue;
                const double *pr = yr.data() + (size_t)k * n, *pi = yi.data() + (size_t)k * n;
                for (int j = 0; j < n; j++) {
                    qr[j] += a * pr[j] - b * pi[j];
                    qi[j] += a * pi[j] + b * pr[j];
                }
            }
        }
        std::vector<cd> Z((size_t)n * n);
        for (size_t t = 0; t < (size_t)n * n; t++) Z[t] = cd(zr[t], zi[t]);
        return Z;
    };
    double norm1 = 0;
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += std::abs(A[(size_t)i * n + j]);
        norm1 = std::max(norm1, s);
    }
    const double theta13 = 5.371920351148152;
    int s = 0;
    if (norm1 > theta13) s = std::max(0, (int)std::ceil(std::log2(norm1 / theta13)));
    double scale = std::ldexp(1.0, -s);
    std::vector<cd> X(A);
    for (auto &v : X) v *= scale;
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0,  129060195264000.0,   10559470521600.0,
                                 670442572800.0,      33522128640.0,       1323241920.0,
                                 40840800.0,          960960.0,            16380.0,
                                 182.0,               1.0};
    std::vector<cd> I((size_t)n * n, cd(0, 0));
    for (int i = 0; i < n; i++) I[(size_t)i * n + i] = 1;
    auto X2 = mm(X, X), X4 = mm(X2, X2), X6 = mm(X4, X2);
    std::vector<cd> U1((size_t)n * n), V1((size_t)n * n), T((size_t)n * n), W((size_t)n * n);
    for (size_t i = 0; i < I.size(); i++) {
        T[i] = b[13] * X6[i] + b[11] * X4[i] + b[9] * X2[i];
        W[i] = b[12] * X6[i] + b[10] * X4[i] + b[8] * X2[i];
    }
    auto T6 = mm(X6, T), W6 = mm(X6, W);
    for (size_t i = 0; i < I.size(); i++) {
        U1[i] = T6[i] + b[7] * X6[i] + b[5] * X4[i] + b[3] * X2[i] + b[1] * I[i];
        V1[i] = W6[i] + b[6] * X6[i] + b[4] * X4[i] + b[2] * X2[i] + b[0] * I[i];
    }
    auto U = mm(X, U1);
    // solve (V - U) E = (V + U)
    std::vector<cd> P((size_t)n * n), Q((size_t)n * n);
    for (size_t i = 0; i < I.size(); i++) {
        P[i] = V1[i] - U[i];
        Q[i] = V1[i] + U[i];
    }
    // Gaussian elimination with partial pivoting on P, applied to Q (Q on
    // split planes: its row updates are the bulk of the work)
    std::vector<double> &qr = yr, &qi = yi;
    for (size_t t = 0; t < (size_t)n * n; t++) {
        qr[t] = Q[t].real();
        qi[t] = Q[t].imag();
    }
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::abs(P[(size_t)r * n + c]) > std::abs(P[(size_t)piv * n + c])) piv = r;
        if (piv != c)
            for (int j = 0; j < n; j++) {
                std::swap(P[(size_t)c * n + j], P[(size_t)piv * n + j]);
                std::swap(qr[(size_t)c * n + j], qr[(size_t)piv * n + j]);
                std::swap(qi[(size_t)c * n + j], qi[(size_t)piv * n + j]);
            }
        const cd d = P[(size_t)c * n + c];
        for (int r = c + 1; r < n; r++) {
            const cd f = P[(size_t)r * n + c] / d;
            if (f == cd(0, 0)) continue;
            for (int j = c; j < n; j++) P[(size_t)r * n + j] -= f * P[(size_t)c * n + j];
            const double a = f.real(), b = f.imag();
            double *rr = qr.data() + (size_t)r * n, *ri = qi.data() + (size_t)r * n;
            const double *cr = qr.data() + (size_t)c * n, *ci = qi.data() + (size_t)c * n;
            for (int j = 0; j < n; j++) {
                rr[j] -= a * cr[j] - b * ci[j];
                ri[j] -= a * ci[j] + b * cr[j];
            }
        }
    }
    for (int c = n - 1; c >= 0; c--) {
        double *rr = qr.data() + (size_t)c * n, *ri = qi.data() + (size_t)c * n;
        for (int k = c + 1; k < n; k++) {
            const double a = P[(size_t)c * n + k].real(), b = P[(size_t)c * n + k].imag();
            const double *kr = qr.data() + (size_t)k * n, *ki = qi.data() + (size_t)k * n;
            for (int j = 0; j < n; j++) {
                rr[j] -= a * kr[j] - b * ki[j];
                ri[j] -= a * ki[j] + b * kr[j];
            }
        }
        const cd dinv = 1.0 / P[(size_t)c * n + c];
        for (int j = 0; j < n; j++) {
            const cd q = cd(rr[j], ri[j]) * dinv;
            rr[j] = q.real();
            ri[j] = q.imag();
        }
    }
    for (size_t t = 0; t < (size_t)n * n; t++) Q[t] = cd(qr[t], qi[t]);
    E = Q;
    for (int i = 0; i < s; i++) E = mm(E, E);
}

namespace {

// Host workers kept for the life of the process, so a lockstep round's
// per-request host work (scalar replay, small exponentials) costs no thread
// creation.  One job at a time: a caller that finds the pool busy (another
// serving thread, a nested call) or finds itself in a forked child runs its
// own threads instead.
class WorkPool {
public:
    static WorkPool &get() {
        static WorkPool *p = new WorkPool;   // never destroyed: workers outlive static teardown
        return *p;
    }
    bool run(size_t n, size_t nt, const std::function<void(size_t)> &fn) {
        if (getpid() != pid_) return false;
        std::unique_lock<std::mutex> own(owner_, std::try_to_lock);
        if (!own.owns_lock()) return false;
        while (workers_.size() + 1 < nt) {
            const size_t id = workers_.size();
            workers_.emplace_back([this, id] { loop(id); });
            workers_.back().detach();
        }
        {
            std::lock_guard<std::mutex> g(m_);
            fn_ = &fn;
            n_ = n;
            next_ = 0;
            err_ = nullptr;
            want_ = nt - 1;
            busy_ = nt - 1;
            gen_++;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [&] { return busy_ == 0; });
        fn_ = nullptr;
        std::exception_ptr e = err_;
        g.unlock();
        if (e) std::rethrow_exception(e);
        return true;
    }

private:
    WorkPool() : pid_(getpid()) {}
    void drain() {
        for (size_t i; (i = next_.fetch_add(1)) < n_;) {
            try {
                (*fn_)(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(m_);
                if (!err_) err_ = std::current_exception();
            }
        }
    }
    void loop(size_t id) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> g(m_);
        for (;;) {
            cv_.wait(g, [&] { return gen_ != seen; });
            seen = gen_;
            if (id >= want_) continue;
            g.unlock();
            drain();
            g.lock();
            if (--busy_ == 0) done_.notify_all();
        }
    }
    const pid_t pid_;
    std::mutex owner_, m_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> workers_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0, want_ = 0, busy_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    std::exception_ptr err_;
};

}  // namespace

void parallel_for(size_t n, const std::function<void(size_t)> &fn) {
    const size_t nt = std::min<size_t>({n, 16, std::max(1u, std::thread::hardware_concurrency())});
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    if (WorkPool::get().run(n, nt, fn)) return;
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex em;
    auto work = [&] {
        for (size_t i; (i = next.fetch_add(1)) < n;) {
            try {
                fn(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(em);
                if (!err) err = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

void stream_pool_setup() {
    // the default pool of the device keeps the memory freed in stream order
    // for reuse (a lockstep batch round re-allocates tens of GB of Krylov
    // bases); stream_pool_trim hands the excess back after large batches
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipMemPool_t pool;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        uint64_t keep = UINT64_MAX;
        if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) (void)hipGetLastError();
    });
}

void stream_pool_trim() {
    int dev = 0;
    hipMemPool_t pool;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    if (hipMemPoolTrimTo(pool, (size_t)8 << 30) != hipSuccess) (void)hipGetLastError();
}

JobTable::~JobTable() {
    if (s_) (void)hipStreamSynchronize(s_);
    if (dev_) (void)hipFreeAsync(dev_, s_);
    if (host_) (void)hipHostFree(host_);
    (void)hipGetLastError();
}

const void *JobTable::stage(const void *src, size_t bytes, hipStream_t s) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (s_ && s != s_) {
        // copies staged on another stream may still read the host buffer
        QS_HIP(hipStreamSynchronize(s_));
        used_ = 0;
    }
    if (used_ + need > cap_) {
        if (s_) QS_HIP(hipStreamSynchronize(s_));
        used_ = 0;
        if (need > cap_) {
            if (dev_) QS_HIP(hipFreeAsync(dev_, s_));
            if (host_) QS_HIP(hipHostFree(host_));
            dev_ = host_ = nullptr;
            cap_ = 0;
            const size_t cap = std::max<size_t>(need, (size_t)16 << 20);
            stream_pool_setup();
            if (hipHostMalloc((void **)&host_, cap, hipHostMallocDefault) != hipSuccess) {
                host_ = nullptr;
                (void)hipGetLastError();
                fail(QS_EOOM, "pinned allocation of " + std::to_string(cap) + " bytes failed");
            }
            if (hipMallocAsync((void **)&dev_, cap, s) != hipSuccess) {
                dev_ = nullptr;
                (void)hipGetLastError();
                fail(QS_EOOM, "device allocation of " + std::to_string(cap) + " bytes failed");
            }
            cap_ = cap;
        }
    }
    s_ = s;
    std::memcpy(host_ + used_, src, bytes);
    QS_HIP(hipMemcpyAsync(dev_ + used_, host_ + used_, bytes, hipMemcpyHostToDevice, s));
    const void *at = dev_ + used_;
    used_ += need;
    return at;
}

namespace {
thread_local hipStream_t t_stream = nullptr;
}

hipStream_t op_stream(const OpHost &op) { return t_stream ? t_stream : op.stream; }
StreamScope::StreamScope(hipStream_t s) : prev(t_stream) { t_stream = s; }
StreamScope::~StreamScope() { t_stream = prev; }

}  // namespace qs
