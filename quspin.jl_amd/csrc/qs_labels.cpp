// qs_labels.cpp -- the reference's representative labels and normalizations.
//
// QuSpin.jl labels each sector column with the parent state of the first
// row holding the largest |P| entry (symmetry_basis.jl:215-232) and reports
// normalization = 1/|P entry|^2 (discrete_basis.jl:2401-2408).  The entries
// come from a stage-by-stage construction (cyclic projector for the first
// symmetry, symmetry_basis.jl:57-108, then P'SP intersections,
// :110-213), and they tie only up to rounding, so the label depends on the
// exact floating-point sequence.  This file repeats that sequence column by
// column: each column's rows are the orbit of its minimum state, so every
// column is computed independently (threads over columns), with Julia's
// complex multiply (no fused multiply-add), its robust complex division
// (Baudin & Smith 2012), LinearAlgebra.generic_norm2, SparseArrays'
// ascending-row accumulation order, and a correctly rounded hypot.
#include <algorithm>
#include <cmath>
#include <thread>

#include "qs_internal.h"

#pragma STDC FP_CONTRACT OFF

namespace qs {

namespace {

struct jc {
    double re, im;
};
inline jc jmul(jc a, jc b) {
    volatile double t1 = a.re * b.re, t2 = a.im * b.im, t3 = a.re * b.im, t4 = a.im * b.re;
    return jc{t1 - t2, t3 + t4};
}
inline jc jadd(jc a, jc b) { return jc{a.re + b.re, a.im + b.im}; }
inline jc jconj(jc a) { return jc{a.re, -a.im}; }

double cdiv2(double a, double b, double c, double d, double r, double t) {
    if (r != 0) {
        volatile double br = b * r;
        if (br != 0) return (a + br) * t;
        volatile double at = a * t, bt = b * t;
        volatile double btr = bt * r;
        return at + btr;
    }
    volatile double bc = b / c;
    volatile double dbc = d * bc;
    return (a + dbc) * t;
}
jc jdiv(jc z, jc w) {
    double a = z.re, b = z.im, c = w.re, d = w.im;
    double ab = std::max(std::fabs(a), std::fabs(b));
    double cd_ = std::max(std::fabs(c), std::fabs(d));
    const double eps = 2.220446049250313e-16;
    const double halfov = 0.5 * 1.7976931348623157e308;
    const double twoun = 2.2250738585072014e-308 * 2.0 / eps;
    const double bs = 2.0 / (eps * eps);
    double s = 1.0;
    if (ab >= halfov) { a *= 0.5; b *= 0.5; s *= 2.0; }
    if (cd_ >= halfov) { c *= 0.5; d *= 0.5; s *= 0.5; }
    if (ab <= twoun) { a *= bs; b *= bs; s /= bs; }
    if (cd_ <= twoun) { c *= bs; d *= bs; s *= bs; }
    double p, q;
    auto cdiv1 = [](double a1, double b1, double c1, double d1, double &p1, double &q1) {
        volatile double r = d1 / c1;
        volatile double dr = d1 * r;
        volatile double t = 1.0 / (c1 + dr);
        p1 = cdiv2(a1, b1, c1, d1, r, t);
        q1 = cdiv2(b1, -a1, c1, d1, r, t);
    };
    if (std::fabs(d) <= std::fabs(c)) {
        cdiv1(a, b, c, d, p, q);
    } else {
        cdiv1(b, a, d, c, p, q);
        q = -q;
    }
    return jc{p * s, q * s};
}

double jhypot(double x, double y) {
    if (std::isinf(x) || std::isinf(y)) return INFINITY;
    double ax = std::fabs(x), ay = std::fabs(y);
    if (ay > ax) std::swap(ax, ay);
    const double eps = 2.220446049250313e-16;
    if (ay <= ax * std::sqrt(eps / 2)) return ax;
    double scale = std::ldexp(1.0, -563);
    if (ax > std::sqrt(1.7976931348623157e308 / 2)) {
        ax *= scale; ay *= scale; scale = 1.0 / scale;
    } else if (ay < std::sqrt(2.2250738585072014e-308)) {
        ax /= scale; ay /= scale;
    } else {
        scale = 1.0;
    }
    double h = std::sqrt(std::fma(ax, ax, ay * ay));
    volatile double hsq = h * h, axsq = ax * ax;
    h -= (std::fma(-ay, ay, hsq - axsq) + std::fma(h, h, -hsq) - std::fma(ax, ax, -axsq)) / (2 * h);
    return h * scale;
}
inline double jabs(jc z) { return jhypot(z.re, z.im); }
inline double jabs2(jc z) {
    volatile double a = z.re * z.re, b = z.im * z.im;
    return a + b;
}
bool japprox(jc x, jc y, double atol, double rtol) {
    if (x.re == y.re && x.im == y.im) return true;
    double d = jabs(jc{x.re - y.re, x.im - y.im});
    double m = std::max(jabs(x), jabs(y));
    return d <= std::max(atol, rtol * m);
}
double gnorm2(const jc *v, int n) {
    double maxabs = 0;
    for (int i = 0; i < n; i++) maxabs = std::max(maxabs, jabs(v[i]));
    if (maxabs == 0 || std::isinf(maxabs)) return maxabs;
    double sum = jabs2(v[0]);
    for (int i = 1; i < n; i++) sum += jabs2(v[i]);
    return std::sqrt(sum);
}

struct StageT {
    int kind;   // 0 translate, 1 parity, 2 flip, 3 parity+flip, 4 flip odd, 5 flip even, 6 general map
    jc lam;
    const GenMap *map = nullptr;
};

uint64_t stage_apply(const StageT &st, uint64_t u, int L, int a, uint64_t odd, uint64_t even) {
    switch (st.kind) {
    case 0: return rotl(u, a, L);
    case 1: return bit_reverse(u, L);
    case 2: return u ^ full_mask(L);
    case 3: return bit_reverse(u, L) ^ full_mask(L);
    case 4: return u ^ odd;
    case 5: return u ^ even;
    case 6: {
        // basis.jl:325-338: the (inverted) bit of each source site moves to its target
        const uint64_t v = u ^ st.map->flip;
        uint64_t r = 0;
        for (int i = 0; i < L; i++)
            if (v >> i & 1) r |= 1ULL << st.map->perm[(size_t)i];
        return r;
    }
    }
    return u;
}

// One column's label and normalization: the orbit O of s0 under the group
// (gapply(e, s)), each stage's map sapply(stage, s).  S: uint64_t, or WState
// for states of up to 256 sites.
template <typename S, typename GA, typename SA>
void emulate_column(int order, const std::vector<StageT> &stages, const S &s0, GA gapply, SA sapply, S &label,
                    double &rnorm) {
    std::vector<S> O;
    for (int e = 0; e < order; e++) O.push_back(gapply(e, s0));
    std::sort(O.begin(), O.end());
    O.erase(std::unique(O.begin(), O.end()), O.end());
    const int R = (int)O.size();
    auto idx = [&](const S &u) { return (int)(std::lower_bound(O.begin(), O.end(), u) - O.begin()); };
    std::vector<jc> val(R);
    std::vector<int> col(R, -1);
    const int DEAD = -2;
    // stage 1: cyclic projector (symmetry_basis.jl:57-108)
    {
        const StageT &st = stages[0];
        std::vector<int> cyc;
        std::vector<jc> coef;
        for (int i = 0; i < R; i++) {
            if (col[i] != -1) continue;
            cyc.assign(1, i);
            int cur = i;
            for (;;) {
                int nx = idx(sapply(st, O[cur]));
                if (nx == i) break;
                cyc.push_back(nx);
                cur = nx;
            }
            int len = (int)cyc.size();
            coef.assign(len, jc{1.0, 0.0});
            for (int q = 1; q < len; q++) coef[q] = jdiv(jmul(coef[q - 1], jc{1.0, 0.0}), st.lam);
            jc closure = jmul(coef[len - 1], jc{1.0, 0.0});
            bool keep = japprox(closure, jmul(st.lam, coef[0]), 2e-10, 2e-10);
            if (!keep) {
                for (int r : cyc) col[r] = DEAD;
                continue;
            }
            double nrm = gnorm2(coef.data(), len);
            for (int q = 0; q < len; q++) {
                val[cyc[q]] = jc{coef[q].re / nrm, coef[q].im / nrm};
                col[cyc[q]] = i;
            }
        }
    }
    // later stages: P'SP intersections (symmetry_basis.jl:110-213): the
    // stage maps each column onto one column with a phase; the cycles of that
    // column map, walked from their first column, give the new columns
    for (size_t si = 1; si < stages.size(); si++) {
        const StageT &st = stages[si];
        std::vector<int> ncol(R, -1);
        std::vector<jc> nval(val);
        std::vector<char> seen(R, 0);
        auto overlap = [&](int from) {
            // R[to, from] = sum over rows r of `from` (ascending) of conj(P[S r, to]) * P[r, from]
            jc acc{0, 0};
            bool first = true;
            for (int r = 0; r < R; r++) {
                if (col[r] != from) continue;
                jc prod = jmul(jconj(val[idx(sapply(st, O[r]))]), val[r]);
                acc = first ? prod : jadd(acc, prod);
                first = false;
            }
            return acc;
        };
        std::vector<int> cyc;
        std::vector<jc> ph, c;
        for (int i = 0; i < R; i++) {
            if (col[i] != i || seen[i]) continue;
            cyc.clear();
            ph.clear();
            bool dead = false;
            int cur = i;
            for (;;) {
                cyc.push_back(cur);
                seen[cur] = 1;
                ph.push_back(overlap(cur));
                const int nx = col[idx(sapply(st, O[cur]))];
                if (nx == i) break;
                if (nx < 0 || seen[nx]) {
                    dead = true;   // no column cycle through i (not reached by a monomial stage)
                    break;
                }
                cur = nx;
            }
            const int len = (int)cyc.size();
            bool keep = !dead;
            c.assign(len, jc{1.0, 0.0});
            if (keep) {
                for (int q = 0; q + 1 < len; q++) c[q + 1] = jdiv(jmul(c[q], ph[q]), st.lam);
                jc closure = jmul(c[len - 1], ph[len - 1]);
                keep = japprox(closure, jmul(st.lam, c[0]), 2e-10, 2e-10);
            }
            const double nrm = gnorm2(c.data(), len);
            for (int q = 0; q < len; q++) {
                const jc w{c[q].re / nrm, c[q].im / nrm};
                for (int r = 0; r < R; r++) {
                    if (col[r] != cyc[q]) continue;
                    ncol[r] = keep ? i : DEAD;
                    if (keep) nval[r] = jmul(val[r], w);
                }
            }
        }
        for (int r = 0; r < R; r++)
            if (col[r] == DEAD) ncol[r] = DEAD;
        col.swap(ncol);
        val.swap(nval);
    }
    // symmetry_basis.jl:221-229: first row (ascending) with the largest |entry|
    int best = -1;
    for (int r = 0; r < R; r++) {
        if (col[r] != 0) continue;
        if (best < 0 || jabs(val[r]) > jabs(val[best])) best = r;
    }
    if (best < 0) {   // cannot happen for a kept column
        label = s0;
        rnorm = (double)R;
        return;
    }
    label = O[best];
    rnorm = 1.0 / jabs2(val[best]);
}

}  // namespace

void compute_reference_labels(BasisHost &b) {
    const int64_t n = b.Ns;
    std::vector<uint64_t> reps((size_t)n);
    if (b.use64) {
        QS_HIP(hipMemcpy(reps.data(), b.reps64.get(), sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    } else {
        std::vector<uint32_t> r32((size_t)n);
        QS_HIP(hipMemcpy(r32.data(), b.reps32.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) reps[i] = r32[i];
    }
    // the reference's stage order: k first, then p, z, pz, zA, zB (basis.jl:704-763);
    // SpinBasisGeneral: its blocks in block order (basis.jl:575-602)
    std::vector<StageT> stages;
    for (const GenMap &m : b.gmaps) {
        const int q = ((m.q % m.order) + m.order) % m.order;
        const double ang = (6.283185307179586 * (double)q) / (double)m.order;
        stages.push_back(StageT{6, jc{std::cos(ang), std::sin(ang)}, &m});
    }
    if (b.blocks[0] != QS_UNSET) {
        int order = b.L / b.a;
        int m = ((b.blocks[0] % order) + order) % order;
        double ang = (6.283185307179586 * (double)m) / (double)order;
        stages.push_back(StageT{0, jc{std::cos(ang), std::sin(ang)}});
    }
    for (int i = 1; i < 6; i++)
        if (b.blocks[i] != QS_UNSET) stages.push_back(StageT{i, jc{(double)b.blocks[i], 0.0}});
    uint64_t odd = 0, even = 0;
    for (int site = 1; site <= b.L; site += 2) odd |= 1ULL << (site - 1);
    for (int site = 2; site <= b.L; site += 2) even |= 1ULL << (site - 1);
    b.labels.resize((size_t)n);
    b.ref_norm.resize((size_t)n);
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; w++)
        pool.emplace_back([&, w] {
            const GroupHost &g = b.group;
            auto ga = [&](int e, uint64_t u) { return group_apply(g, e, u); };
            auto sa = [&](const StageT &st, uint64_t u) { return stage_apply(st, u, b.L, b.a, odd, even); };
            for (int64_t i = w; i < n; i += nt)
                emulate_column(g.order(), stages, reps[i], ga, sa, b.labels[i], b.ref_norm[i]);
        });
    for (auto &t : pool) t.join();
    b.labels_ready = true;
}

// Labels and normalizations of a wide basis: the same stage sequence on
// states of up to 256 sites (general maps only).
void compute_wide_labels(BasisHost &b) {
    const int64_t n = b.Ns;
    std::vector<uint64_t> words((size_t)4 * n);
    QS_HIP(hipMemcpy(words.data(), b.wreps.get(), sizeof(uint64_t) * words.size(), hipMemcpyDeviceToHost));
    std::vector<StageT> stages;
    for (const GenMap &m : b.gmaps) {
        const int q = ((m.q % m.order) + m.order) % m.order;
        const double ang = (6.283185307179586 * (double)q) / (double)m.order;
        stages.push_back(StageT{6, jc{std::cos(ang), std::sin(ang)}, &m});
    }
    const int L = b.L;
    auto sa = [L](const StageT &st, const WState &u) {
        // basis.jl:325-338: the (inverted) bit of each source site moves to its target
        const WState v = u ^ st.map->flipw;
        WState r;
        for (int i = 0; i < L; i++)
            if (v.bit(i)) r.set(st.map->perm[(size_t)i]);
        return r;
    };
    const GroupHost &g = b.group;
    auto ga = [&g](int e, const WState &u) { return wide_group_apply(g, e, u); };
    b.wlabels.assign((size_t)n, WState());
    b.ref_norm.assign((size_t)n, 0.0);
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; w++)
        pool.emplace_back([&, w] {
            for (int64_t i = w; i < n; i += nt) {
                WState s0;
                for (int k = 0; k < 4; k++) s0.w[k] = words[(size_t)4 * i + k];
                if (stages.empty()) {
                    b.wlabels[(size_t)i] = s0;
                    b.ref_norm[(size_t)i] = 1.0;
                    continue;
                }
                emulate_column(g.order(), stages, s0, ga, sa, b.wlabels[(size_t)i], b.ref_norm[(size_t)i]);
            }
        });
    for (auto &t : pool) t.join();
    b.labels.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) b.labels[(size_t)i] = b.wlabels[(size_t)i].w[0];
    b.labels_ready = true;
}

}  // namespace qs
