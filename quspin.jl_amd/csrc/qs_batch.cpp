// qs_batch.cpp -- many sectors / initial states / time grids in lockstep.
//
// The reference evolves the blocks of a BlockOps one after another (or on
// host threads): per block P'psi, _krylov_expmv_times, then P * states
// (tools.jl:2450-2560, operators.jl:1655-1893).  Here every block's Krylov
// recursion advances together: one step of all blocks is one batched matvec
// launch plus batched Gram / update / norm launches and a single host round
// trip for all their scalars.  Each job inside a batched launch sees the
// block partition of its own single launch, and the host arithmetic is the
// single driver's (KrylovProj), so a job's result does not depend on which
// other jobs share the batch: alone and in a mix it is bitwise the same.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "qs_internal.h"

namespace qs {

namespace {

// ---------------------------------------------------- batched vector ops --
struct BWork {
    explicit BWork(hipStream_t st) : s(st) {}
    hipStream_t s;   // every launch using this workspace runs on it
    SBuf<double> partial, result, coef;
    std::vector<double> host;
    JobTable jt;     // job tables of launches over many requests
    void need(size_t p, size_t r, size_t c) {
        if (partial.n < p) partial.alloc(std::max<size_t>(p, 4096), s);
        if (result.n < r) result.alloc(std::max<size_t>(r, 1024), s);
        if (coef.n < c) coef.alloc(std::max<size_t>(c, 1024), s);
    }
};

struct NormReq {
    const void *x;
    int64_t n;
};

void bnorm2(int dtype, const std::vector<NormReq> &rq, BWork &wk, hipStream_t s, std::vector<double> &nrm2) {
    size_t np = 0;
    std::vector<size_t> po;
    for (auto &q : rq) {
        po.push_back(np);
        np += (size_t)2 * vec_blocks(q.n);
    }
    wk.need(np, (size_t)2 * rq.size(), 0);
    std::vector<VecJob> a, r;
    std::vector<int> ab, rb;
    for (size_t i = 0; i < rq.size(); i++) {
        VecJob j{};
        j.x = rq[i].x;
        j.n = rq[i].n;
        j.m = 1;
        j.nb = vec_blocks(rq[i].n);
        j.partial = wk.partial.get() + po[i];
        a.push_back(j);
        ab.push_back(j.nb);
        j.aux = wk.result.get() + 2 * i;
        r.push_back(j);
        rb.push_back(1);
    }
    launch_vec_jobs(dtype, VOP_NORM2, a, ab, s, &wk.jt);
    launch_vec_jobs(dtype, VOP_REDUCE, r, rb, s, &wk.jt);
    wk.host.resize(2 * rq.size());
    QS_HIP(hipMemcpyAsync(wk.host.data(), wk.result.get(), sizeof(double) * 2 * rq.size(), hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    nrm2.resize(rq.size());
    for (size_t i = 0; i < rq.size(); i++) nrm2[i] = wk.host[2 * i];
}

struct AxpbyReq {          // y = a x + b y
    cd a;
    const void *x;
    cd b;
    void *y;
    int64_t n;
};

void baxpby(int dtype, const std::vector<AxpbyReq> &rq, hipStream_t s, JobTable *jt = nullptr) {
    std::vector<VecJob> a;
    std::vector<int> ab;
    for (auto &q : rq) {
        VecJob j{};
        j.x = q.x;
        j.w = q.y;
        j.n = q.n;
        j.ar = q.a.real();
        j.ai = q.a.imag();
        j.br = q.b.real();
        j.bi = q.b.imag();
        j.b_zero = q.b == cd(0, 0);
        j.nb = ew_blocks(q.n);
        a.push_back(j);
        ab.push_back(j.nb);
    }
    launch_vec_jobs(dtype, VOP_AXPBY, a, ab, s, jt);
}

struct CombReq {           // Y[:, :k] = V[:, :m] C (C column-major m x k)
    const void *V;
    int64_t n;
    int m;
    std::vector<cd> C;
    int k;
    void *Y;
};

void bcombine(int dtype, const std::vector<CombReq> &rq, BWork &wk, hipStream_t s) {
    size_t nc = 0;
    std::vector<size_t> co;
    for (auto &q : rq) {
        co.push_back(nc);
        nc += (size_t)2 * q.m * q.k;
    }
    wk.need(0, 0, nc);
    std::vector<double> cc(nc);
    for (size_t i = 0; i < rq.size(); i++)
        for (size_t t = 0; t < (size_t)rq[i].m * rq[i].k; t++) {
            cc[co[i] + 2 * t] = rq[i].C[t].real();
            cc[co[i] + 2 * t + 1] = rq[i].C[t].imag();
        }
    // staged through the pinned job table: no wait before the next call
    const double *coef = nc ? (const double *)wk.jt.stage(cc.data(), sizeof(double) * nc, s) : nullptr;
    std::vector<VecJob> a;
    std::vector<int> ab;
    for (size_t i = 0; i < rq.size(); i++) {
        // a job with many output columns is a GEMM: its own launch on the matrix cores
        if (combine_on_mfma(dtype, rq[i].V, rq[i].n, rq[i].m, coef + co[i], rq[i].k, rq[i].Y, s)) continue;
        VecJob j{};
        j.V = rq[i].V;
        j.n = rq[i].n;
        j.m = rq[i].m;
        j.k = rq[i].k;
        j.w = rq[i].Y;
        j.aux = (double *)coef + co[i];
        j.nb = ew_blocks(rq[i].n);
        a.push_back(j);
        ab.push_back(j.nb);
    }
    if (!a.empty()) launch_vec_jobs(dtype, VOP_COMBINE, a, ab, s, &wk.jt);
}

// Device-scalar forms: results land in device memory at the caller's
// addresses and coefficients are read from there, so a whole lockstep
// factorization is queued without a host round trip.  `wk.partial` must
// already hold every job's partials (bwork_reserve).
void bwork_reserve(BWork &wk, const std::vector<int64_t> &ns, int mmax) {
    size_t np = 0;
    for (int64_t n : ns) np += (size_t)2 * vec_blocks(n) * std::max(1, mmax);
    wk.need(np, 0, 0);
}

struct DotDev {            // out[2i..2i+1] = <V[:, i], w>, i < m
    const void *V;
    int64_t n;
    int m;
    const void *w;
    double *out;
};

void bdots_dev(int dtype, const std::vector<DotDev> &rq, BWork &wk, hipStream_t s) {
    std::vector<VecJob> g, r;
    std::vector<int> gb, rb;
    size_t np = 0;
    for (auto &q : rq) {
        VecJob j{};
        j.V = q.V;
        j.x = q.w;
        j.n = q.n;
        j.m = q.m;
        j.nb = vec_blocks(q.n);
        j.partial = wk.partial.get() + np;
        j.aux = q.out;
        np += (size_t)2 * j.nb * q.m;
        g.push_back(j);
        gb.push_back(j.nb);
        r.push_back(j);
        rb.push_back(j.m);
    }
    if (np > wk.partial.n) fail(QS_EARG, "internal: batched workspace not reserved");
    launch_vec_jobs(dtype, VOP_GRAM, g, gb, s, &wk.jt);
    launch_vec_jobs(dtype, VOP_REDUCE, r, rb, s, &wk.jt);
}

struct UpdDev {            // w -= V[:, :m] h (h device); nrm2 (device, optional) = ||w||^2
    const void *V;
    int64_t n;
    int m;
    const double *h;
    void *w;
    double *nrm2;
};

void bupdate_dev(int dtype, const std::vector<UpdDev> &rq, BWork &wk, hipStream_t s) {
    std::vector<VecJob> u, r;
    std::vector<int> ub, rb;
    size_t np = 0;
    for (auto &q : rq) {
        VecJob j{};
        j.V = q.V;
        j.w = q.w;
        j.n = q.n;
        j.m = q.m;
        j.nb = vec_blocks(q.n);
        j.partial = wk.partial.get() + np;
        j.aux = const_cast<double *>(q.h);
        np += (size_t)2 * j.nb;
        u.push_back(j);
        ub.push_back(j.nb);
        if (q.nrm2) {
            VecJob t = j;
            t.aux = q.nrm2;
            r.push_back(t);
            rb.push_back(1);
        }
    }
    if (np > wk.partial.n) fail(QS_EARG, "internal: batched workspace not reserved");
    launch_vec_jobs(dtype, VOP_UPDATE, u, ub, s, &wk.jt);
    if (!r.empty()) launch_vec_jobs(dtype, VOP_REDUCE, r, rb, s, &wk.jt);
}

void bnorm2_dev(int dtype, const std::vector<std::pair<const void *, int64_t>> &xs, const std::vector<double *> &out,
                BWork &wk, hipStream_t s) {
    std::vector<VecJob> a, r;
    std::vector<int> ab, rb;
    size_t np = 0;
    for (size_t i = 0; i < xs.size(); i++) {
        VecJob j{};
        j.x = xs[i].first;
        j.n = xs[i].second;
        j.m = 1;
        j.nb = vec_blocks(j.n);
        j.partial = wk.partial.get() + np;
        np += (size_t)2 * j.nb;
        a.push_back(j);
        ab.push_back(j.nb);
        j.aux = out[i];
        r.push_back(j);
        rb.push_back(1);
    }
    if (np > wk.partial.n) fail(QS_EARG, "internal: batched workspace not reserved");
    launch_vec_jobs(dtype, VOP_NORM2, a, ab, s, &wk.jt);
    launch_vec_jobs(dtype, VOP_REDUCE, r, rb, s, &wk.jt);
}

struct AxpbyDev {          // y = a x + b y with a from device memory (CoefMode)
    const double *coef;
    int cmode;
    const void *x;
    cd b;
    void *y;
    int64_t n;
};

void baxpby_dev(int dtype, const std::vector<AxpbyDev> &rq, hipStream_t s, JobTable *jt) {
    std::vector<VecJob> a;
    std::vector<int> ab;
    for (auto &q : rq) {
        VecJob j{};
        j.x = q.x;
        j.w = q.y;
        j.n = q.n;
        j.br = q.b.real();
        j.bi = q.b.imag();
        j.b_zero = q.b == cd(0, 0);
        j.acoef = q.coef;
        j.cmode = q.cmode;
        j.nb = ew_blocks(q.n);
        a.push_back(j);
        ab.push_back(j.nb);
    }
    launch_vec_jobs(dtype, VOP_AXPBY, a, ab, s, jt);
}

// ------------------------------------------------ lockstep factorization --
struct BFact {
    int op;
    int64_t n;
    // n x (steps + 1), a slice of one allocation shared by every request of
    // the call (thousands of separate stream-ordered frees cost more host
    // time than the factorizations take on the device)
    std::shared_ptr<SBuf<char>> store;
    char *V = nullptr;
    KrylovProj P;
    void *col(int j, size_t esz) const { return (void *)(V + (size_t)j * n * esz); }
};

// operators.jl:1655-1731 for every request at once; identical per-request
// operation order to the single driver.  All steps of all requests are
// queued with their scalars in device memory (FactScalars layout per
// request); one copy at the end feeds KrylovProj::replay, which applies the
// reference's bookkeeping and breakdown test in step order.
// Krylov bases of a lockstep call in one device allocation reused by every
// round: slot i holds request i's n x (steps + 1) columns (stream order keeps
// a round's reads before the next round's writes).
struct Arena {
    std::shared_ptr<SBuf<char>> buf;
    size_t per = 0, slots = 0;
};

void bfactorize(BatchMatvec &bm, const std::vector<const OpHost *> &ops, int dtype,
                const std::vector<std::pair<int, const void *>> &req, int krylov_dim, BWork &wk, hipStream_t s,
                std::vector<std::unique_ptr<BFact>> &F, std::vector<int> &nmv, const Arena *arena = nullptr) {
    const size_t esz = dtype_size(dtype);
    const size_t R = req.size();
    F.clear();
    std::vector<int64_t> ns;
    std::vector<size_t> off;
    size_t tot = 0;
    int max_steps = 0;
    for (size_t i = 0; i < R; i++) {
        auto f = std::make_unique<BFact>();
        const OpHost &op = *ops[req[i].first];
        f->op = req[i].first;
        f->n = op.basis->Ns;
        f->P.dim = f->n;
        f->P.hermitian = lanczos_ok(op, dtype);
        f->P.steps = (int)std::min<int64_t>(krylov_dim, f->n);
        f->P.H.assign((size_t)(f->P.steps + 1) * f->P.steps, cd(0, 0));
        f->P.completed = f->P.steps;
        f->P.terminal = 0;
        ns.push_back(f->n);
        off.push_back(tot);
        tot += FactScalars(f->P.steps).total();
        max_steps = std::max(max_steps, f->P.steps);
        F.push_back(std::move(f));
    }
    bool in_arena = arena && arena->buf && R <= arena->slots;
    if (in_arena)
        for (auto &f : F) in_arena &= (size_t)f->n * (f->P.steps + 1) * esz <= arena->per;
    if (in_arena) {
        for (size_t i = 0; i < R; i++) {
            F[i]->store = arena->buf;
            F[i]->V = arena->buf->get() + i * arena->per;
        }
    } else {
        size_t bytes = 0;
        for (auto &f : F) bytes += ((size_t)f->n * (f->P.steps + 1) * esz + 255) & ~(size_t)255;
        auto store = std::make_shared<SBuf<char>>(std::max<size_t>(bytes, 1), s);
        size_t at = 0;
        for (auto &f : F) {
            f->store = store;
            f->V = store->get() + at;
            at += ((size_t)f->n * (f->P.steps + 1) * esz + 255) & ~(size_t)255;
        }
    }
    bwork_reserve(wk, ns, max_steps + 1);
    SBuf<double> d(std::max<size_t>(tot, 1), s);
    auto sc = [&](size_t i) { return FactScalars(F[i]->P.steps); };
    {
        std::vector<std::pair<const void *, int64_t>> xs;
        std::vector<double *> outs;
        std::vector<AxpbyDev> ax;
        for (size_t i = 0; i < R; i++) {
            xs.push_back({req[i].second, F[i]->n});
            outs.push_back(d.get() + off[i] + sc(i).beta2());
            ax.push_back({outs.back(), CO_INV_SQRT, req[i].second, cd(0, 0), F[i]->col(0, esz), F[i]->n});
        }
        bnorm2_dev(dtype, xs, outs, wk, s);
        baxpby_dev(dtype, ax, s, &wk.jt);
    }
    for (int col = 0; col < max_steps; col++) {
        std::vector<size_t> act;
        for (size_t i = 0; i < R; i++)
            if (col < F[i]->P.steps) act.push_back(i);
        if (act.empty()) break;
        std::vector<size_t> herm, arn;
        for (size_t i : act) (F[i]->P.hermitian ? herm : arn).push_back(i);
        auto at = [&](size_t i, size_t o) { return d.get() + off[i] + o; };
        // assembled Hermitian requests: the product also subtracts the
        // previous basis vector and leaves the partials of <v_col, w>
        bool fused = !herm.empty();
        for (size_t i : herm) fused = fused && ops[F[i]->op]->sell;
        size_t np = 0;
        if (fused)
            for (size_t i : herm) np += (size_t)2 * ((F[i]->n + 255) / 256);
        wk.need(np, 0, 0);   // (grows the partials in stream order before any pointer into them is taken)
        std::vector<MatvecJob> mj;
        std::vector<VecJob> fr;
        std::vector<int> frb;
        np = 0;
        for (size_t i : act) {
            MatvecJob j{F[i]->op, F[i]->col(col, esz), F[i]->col(col + 1, esz)};
            if (fused && F[i]->P.hermitian) {
                if (col > 0) {
                    j.z = F[i]->col(col - 1, esz);
                    j.zc = at(i, sc(i).nrm2(col - 1));
                    j.zmode = CO_NEG_SQRT;
                }
                j.dot = wk.partial.get() + np;
                VecJob r{};
                r.partial = j.dot;
                r.nb = (int)((F[i]->n + 255) / 256);
                r.aux = at(i, sc(i).h1(col));
                np += (size_t)2 * r.nb;
                fr.push_back(r);
                frb.push_back(1);
            }
            mj.push_back(j);
        }
        launch_batch_matvec(bm, mj, s);
        if (fused) {
            launch_vec_jobs(dtype, VOP_REDUCE, fr, frb, s, &wk.jt);
            std::vector<UpdDev> uq;
            for (size_t i : herm)
                uq.push_back({F[i]->col(col, esz), F[i]->n, 1, at(i, sc(i).h1(col)), F[i]->col(col + 1, esz),
                              at(i, sc(i).nrm2(col))});
            bupdate_dev(dtype, uq, wk, s);
        } else if (!herm.empty()) {
            // three-term Lanczos without reorthogonalisation
            std::vector<AxpbyDev> ax;
            for (size_t i : herm)
                if (col > 0)
                    ax.push_back({at(i, sc(i).nrm2(col - 1)), CO_NEG_SQRT, F[i]->col(col - 1, esz), cd(1, 0),
                                  F[i]->col(col + 1, esz), F[i]->n});
            if (!ax.empty()) baxpby_dev(dtype, ax, s, &wk.jt);
            std::vector<DotDev> dq;
            std::vector<UpdDev> uq;
            for (size_t i : herm) {
                dq.push_back({F[i]->col(col, esz), F[i]->n, 1, F[i]->col(col + 1, esz), at(i, sc(i).h1(col))});
                uq.push_back({F[i]->col(col, esz), F[i]->n, 1, at(i, sc(i).h1(col)), F[i]->col(col + 1, esz),
                              at(i, sc(i).nrm2(col))});
            }
            bdots_dev(dtype, dq, wk, s);
            bupdate_dev(dtype, uq, wk, s);
        }
        if (!arn.empty()) {
            // Arnoldi, two Gram-Schmidt passes
            std::vector<DotDev> d1, d2;
            std::vector<UpdDev> u1, u2;
            for (size_t i : arn) {
                void *w = F[i]->col(col + 1, esz);
                d1.push_back({F[i]->V, F[i]->n, col + 1, w, at(i, sc(i).h1(col))});
                d2.push_back({F[i]->V, F[i]->n, col + 1, w, at(i, sc(i).h2(col))});
                u1.push_back({F[i]->V, F[i]->n, col + 1, at(i, sc(i).h1(col)), w, nullptr});
                u2.push_back({F[i]->V, F[i]->n, col + 1, at(i, sc(i).h2(col)), w, at(i, sc(i).nrm2(col))});
            }
            bdots_dev(dtype, d1, wk, s);
            // first update and second Gram product in one read of the basis
            std::vector<VecJob> ug, rg;
            std::vector<int> rgb;
            size_t np = 0;
            for (size_t i : arn) {
                VecJob j{};
                j.V = F[i]->V;
                j.w = F[i]->col(col + 1, esz);
                j.n = F[i]->n;
                j.m = col + 1;
                j.nb = vec_blocks(F[i]->n);
                j.aux = at(i, sc(i).h1(col));
                j.partial = wk.partial.get() + np;
                np += (size_t)2 * j.nb * (col + 2);
                ug.push_back(j);
                VecJob r = j;
                r.aux = at(i, sc(i).h2(col));   // h2[0 .. m) and |w'|^2 after them (unused)
                rg.push_back(r);
                rgb.push_back(col + 2);
            }
            if (np <= wk.partial.n && launch_updgram_jobs(dtype, ug, s, &wk.jt)) {
                launch_vec_jobs(dtype, VOP_REDUCE, rg, rgb, s, &wk.jt);
            } else {
                bupdate_dev(dtype, u1, wk, s);
                bdots_dev(dtype, d2, wk, s);
            }
            bupdate_dev(dtype, u2, wk, s);
        }
        std::vector<AxpbyDev> sc2;
        for (size_t i : act)
            if (col + 1 < F[i]->P.steps)
                sc2.push_back({at(i, sc(i).nrm2(col)), CO_INV_SQRT, F[i]->col(col + 1, esz), cd(0, 0),
                               F[i]->col(col + 1, esz), F[i]->n});
        if (!sc2.empty()) baxpby_dev(dtype, sc2, s, &wk.jt);
    }
    std::vector<double> host(tot);
    if (tot) QS_HIP(hipMemcpyAsync(host.data(), d.get(), sizeof(double) * tot, hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    std::vector<int> done(R, 0);
    parallel_for(R, [&](size_t i) {
        std::vector<double> mine(host.begin() + off[i], host.begin() + off[i] + sc(i).total());
        done[i] = F[i]->P.replay(sc(i), mine);
    });
    for (size_t i = 0; i < R; i++) nmv[i] += done[i];
}

// One pending _krylov_expmv_vector call (operators.jl:1758-1803).  A call
// owns the split midpoints it reads and writes (vbuf / obuf; null for the
// task's own v and output column) so they live until every call using them
// has run.
struct Call {
    const void *v;
    cd scale;
    double tol;
    int depth;
    void *out;
    std::shared_ptr<char> vbuf, obuf;
    double v2 = -1;   // |v|^2 when known on the host (the previous call's output norm)
};

// Split midpoints of one call of expm_times_batch: fixed-size device blocks
// carved from a few growing chunks and recycled through a free list, so the
// thousands of halvings cost no allocation each (every use is on the one
// stream, whose order keeps a recycled block's earlier reads first).
class MidPool {
public:
    MidPool(size_t bytes, hipStream_t s) : bytes_((bytes + 255) & ~(size_t)255), s_(s) {}
    std::shared_ptr<char> get() {
        if (free_.empty()) {
            const size_t k = std::max<size_t>(16, total_);
            chunks_.push_back(std::make_unique<SBuf<char>>(k * bytes_, s_));
            for (size_t i = 0; i < k; i++) free_.push_back(chunks_.back()->get() + i * bytes_);
            total_ += k;
        }
        char *p = free_.back();
        free_.pop_back();
        return std::shared_ptr<char>(p, [this](char *q) { free_.push_back(q); });
    }
private:
    size_t bytes_;
    hipStream_t s_;
    size_t total_ = 0;
    std::vector<std::unique_ptr<SBuf<char>>> chunks_;
    std::vector<char *> free_;
};
struct Chain {              // the calls of one (task, time point), run in order
    int task;
    std::vector<Call> stack;  // back() runs next
};

}  // namespace

void expm_times_batch(const std::vector<const OpHost *> &ops, int dtype, const std::vector<ExpmTask> &tasks,
                      double tol, int krylov_dim, hipStream_t s, std::vector<int> &nmatvec) {
    if (!dtype_complex(dtype)) fail(QS_EARG, "expm_multiply needs a complex vector dtype");
    if (tol <= 0) fail(QS_EARG, "tol must be positive");
    if (krylov_dim <= 0) fail(QS_EARG, "krylov_dim must be positive");
    const size_t esz = dtype_size(dtype);
    auto bm = make_batch_matvec(ops, dtype);
    BWork wk(s);
    nmatvec.assign(tasks.size(), 0);
    auto outcol = [&](size_t t, size_t j) {
        return (void *)((char *)tasks[t].out + j * ops[tasks[t].op]->basis->Ns * esz);
    };
    auto nrows = [&](size_t t) { return ops[tasks[t].op]->basis->Ns; };
    // tasks with no work: copies of v
    std::vector<NormReq> nq;
    for (size_t t = 0; t < tasks.size(); t++) nq.push_back({tasks[t].v, nrows(t)});
    std::vector<double> v2;
    if (!nq.empty()) bnorm2(dtype, nq, wk, s, v2);
    std::vector<size_t> live;
    for (size_t t = 0; t < tasks.size(); t++) {
        bool all_zero = true;
        for (auto &z : tasks[t].scales) all_zero &= z == cd(0, 0);
        if (tasks[t].scales.empty()) continue;
        if (all_zero || v2[t] == 0) {
            for (size_t j = 0; j < tasks[t].scales.size(); j++)
                QS_HIP(hipMemcpyAsync(outcol(t, j), tasks[t].v, nrows(t) * esz, hipMemcpyDeviceToDevice, s));
            continue;
        }
        live.push_back(t);
    }
    // one factorization per task, all in lockstep
    std::vector<std::pair<int, const void *>> req;
    for (size_t t : live) req.push_back({tasks[t].op, tasks[t].v});
    int64_t nmax = 1;
    for (auto *op : ops) nmax = std::max<int64_t>(nmax, op->basis->Ns);
    Arena arena;
    arena.per = ((size_t)nmax * (krylov_dim + 1) * esz + 255) & ~(size_t)255;
    auto reserve = [&](size_t slots) {
        if (slots <= arena.slots) return;
        arena.buf.reset();
        arena.buf = std::make_shared<SBuf<char>>(slots * arena.per, s);
        arena.slots = slots;
    };
    reserve(req.size());
    MidPool mids((size_t)nmax * esz, s);   // before the chains that hold its blocks
    std::vector<std::unique_ptr<BFact>> F;
    std::vector<int> nmv(req.size(), 0);
    bfactorize(*bm, ops, dtype, req, krylov_dim, wk, s, F, nmv, &arena);
    for (size_t i = 0; i < live.size(); i++) nmatvec[live[i]] += nmv[i];
    // every time point from its task's factorization; failures become chains
    std::vector<Chain> chains;
    std::vector<CombReq> cq;
    std::vector<std::vector<double>> ests(live.size());
    // the tasks' projected problems on host threads, then the combines in task order
    std::vector<std::vector<cd>> tcoefs(live.size());
    std::vector<std::vector<std::vector<cd>>> jcoefs(live.size());
    parallel_for(live.size(), [&](size_t i) {
        const ExpmTask &T = tasks[live[i]];
        const KrylovProj &P = F[i]->P;
        if (P.hermitian) {
            P.lanczos_times(T.scales, tcoefs[i], ests[i]);
        } else {
            ests[i].assign(T.scales.size(), 0.0);
            jcoefs[i].resize(T.scales.size());
            for (size_t j = 0; j < T.scales.size(); j++)
                if (T.scales[j] != cd(0, 0)) ests[i][j] = P.expm_column(T.scales[j], jcoefs[i][j]);
        }
    });
    for (size_t i = 0; i < live.size(); i++) {
        const size_t t = live[i];
        const ExpmTask &T = tasks[t];
        const KrylovProj &P = F[i]->P;
        const int c = P.completed;
        if (P.hermitian) {
            cq.push_back({F[i]->V, F[i]->n, c, std::move(tcoefs[i]), (int)T.scales.size(), T.out});
        } else {
            for (size_t j = 0; j < T.scales.size(); j++) {
                if (T.scales[j] == cd(0, 0)) continue;
                cq.push_back({F[i]->V, F[i]->n, c, std::move(jcoefs[i][j]), 1, outcol(t, j)});
            }
        }
    }
    bcombine(dtype, cq, wk, s);
    std::vector<NormReq> oq;
    std::vector<std::pair<size_t, size_t>> idx;
    for (size_t i = 0; i < live.size(); i++) {
        const size_t t = live[i];
        for (size_t j = 0; j < tasks[t].scales.size(); j++) {
            if (tasks[t].scales[j] == cd(0, 0)) {
                QS_HIP(hipMemcpyAsync(outcol(t, j), tasks[t].v, nrows(t) * esz, hipMemcpyDeviceToDevice, s));
                continue;
            }
            oq.push_back({outcol(t, j), nrows(t)});
            idx.push_back({i, j});
        }
    }
    std::vector<double> o2;
    if (!oq.empty()) bnorm2(dtype, oq, wk, s, o2);
    for (size_t q = 0; q < idx.size(); q++) {
        const size_t i = idx[q].first, j = idx[q].second, t = live[i];
        if (ests[i][j] > tol * std::max(1.0, std::sqrt(o2[q])))
            chains.push_back(
                {(int)t, {Call{tasks[t].v, tasks[t].scales[j], tol, 0, outcol(t, j), nullptr, nullptr, v2[t]}}});
    }
    F.clear();
    // halving recursion, the next call of as many chains as device memory
    // holds Krylov bases for, in lockstep
    size_t fr = 0, tot = 0;
    QS_HIP(hipMemGetInfo(&fr, &tot));
    const size_t per_chain = (size_t)(krylov_dim + 3) * nmax * esz;
    const size_t cap = std::max<size_t>(1, std::min<size_t>(4096, fr / 3 / per_chain));
    if (!chains.empty()) reserve(std::min(cap, chains.size()));
    while (true) {
        std::vector<size_t> run;
        for (size_t c = 0; c < chains.size() && run.size() < cap; c++)
            if (!chains[c].stack.empty()) run.push_back(c);
        if (run.empty()) break;
        // input norms: known on the host for every call but the second half
        // of a split whose first half finished in the same round
        std::vector<NormReq> vq;
        std::vector<size_t> unk;
        std::vector<double> vn(run.size());
        for (size_t r = 0; r < run.size(); r++) {
            const Call &cl = chains[run[r]].stack.back();
            vn[r] = cl.v2;
            if (cl.v2 < 0) {
                vq.push_back({cl.v, nrows(chains[run[r]].task)});
                unk.push_back(r);
            }
        }
        if (!vq.empty()) {
            std::vector<double> got;
            bnorm2(dtype, vq, wk, s, got);
            for (size_t q = 0; q < unk.size(); q++) vn[unk[q]] = got[q];
        }
        std::vector<size_t> fac;
        std::vector<double> fv2;
        for (size_t r = 0; r < run.size(); r++) {
            Chain &ch = chains[run[r]];
            Call &cl = ch.stack.back();
            if (cl.scale == cd(0, 0) || vn[r] == 0) {
                QS_HIP(hipMemcpyAsync(cl.out, cl.v, nrows(ch.task) * esz, hipMemcpyDeviceToDevice, s));
                void *out = cl.out;
                ch.stack.pop_back();
                if (!ch.stack.empty() && ch.stack.back().v == out) ch.stack.back().v2 = vn[r];
                continue;
            }
            fac.push_back(run[r]);
            fv2.push_back(vn[r]);
        }
        if (fac.empty()) continue;
        std::vector<std::pair<int, const void *>> rq;
        for (size_t c : fac) rq.push_back({tasks[chains[c].task].op, chains[c].stack.back().v});
        std::vector<int> m2(rq.size(), 0);
        bfactorize(*bm, ops, dtype, rq, krylov_dim, wk, s, F, m2, &arena);
        std::vector<CombReq> pq;
        std::vector<double> est(fac.size());
        std::vector<std::vector<cd>> coefs(fac.size());
        // the calls' small exponentials are independent: host threads
        parallel_for(fac.size(), [&](size_t r) {
            est[r] = F[r]->P.expm_column(chains[fac[r]].stack.back().scale, coefs[r]);
        });
        for (size_t r = 0; r < fac.size(); r++) {
            nmatvec[chains[fac[r]].task] += m2[r];
            const Call &cl = chains[fac[r]].stack.back();
            pq.push_back({F[r]->V, F[r]->n, F[r]->P.completed, std::move(coefs[r]), 1, cl.out});
        }
        bcombine(dtype, pq, wk, s);
        std::vector<NormReq> on;
        for (size_t r = 0; r < fac.size(); r++) on.push_back({chains[fac[r]].stack.back().out, F[r]->n});
        std::vector<double> onr;
        bnorm2(dtype, on, wk, s, onr);
        F.clear();
        for (size_t r = 0; r < fac.size(); r++) {
            Chain &ch = chains[fac[r]];
            Call cl = ch.stack.back();
            ch.stack.pop_back();
            if (est[r] > cl.tol * std::max(1.0, std::sqrt(onr[r]))) {
                if (cl.depth >= 14) fail(QS_ENOCONV, "Krylov exponential action did not converge; increase krylov_dim");
                std::shared_ptr<char> mid = mids.get();
                // second half (mid -> out) runs after the first (v -> mid): pushed first
                ch.stack.push_back(Call{mid.get(), cl.scale / 2.0, cl.tol / 2, cl.depth + 1, cl.out, mid, cl.obuf});
                ch.stack.push_back(
                    Call{cl.v, cl.scale / 2.0, cl.tol / 2, cl.depth + 1, mid.get(), cl.vbuf, mid, fv2[r]});
            } else if (!ch.stack.empty() && ch.stack.back().v == cl.out) {
                ch.stack.back().v2 = onr[r];   // the next call reads this output
            }
        }
    }
    QS_HIP(hipStreamSynchronize(s));
    stream_pool_trim();
}

}  // namespace qs
