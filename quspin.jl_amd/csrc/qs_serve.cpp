// qs_serve.cpp -- request loop for eigsh / evolve jobs on one GPU.
//
// The reference answers one call at a time on the host.  A serving process
// instead receives eigsh (operators.jl:3367-3408) and evolve /
// _krylov_expmv_times (operators.jl:1805-1893, 2073-2192) requests while
// others run.  Here a fixed set of workers, each with its own non-blocking
// HIP stream, takes requests first-in first-out; a request is admitted when
// a worker is free and its Krylov working set fits the HBM budget beside the
// running ones, so new requests start as soon as slots and memory free up
// and their kernels overlap on the device with the running requests'.
// Every kernel reduces in a fixed order with no atomics, so a request's
// result is bitwise the same alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "qs_internal.h"

struct qs_op {
    qs::OpHost o;
};

namespace qs {

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct ServeJob {
    int64_t id = 0;
    qs_job spec{};
    OpHost *op = nullptr;
    std::vector<char> v0;
    std::vector<double> scales;
    size_t need = 0;
    std::vector<double> vals;
    std::vector<char> out;
    int nconv = 0, nmv = 0;
    int status = QS_OK;
    std::string err;
    double t_submit = 0, t_start = 0, t_end = 0;
    bool done = false;
};

}  // namespace

struct Server {
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeJob *> queue;
    std::unordered_map<int64_t, std::unique_ptr<ServeJob>> jobs;
    size_t budget = 0, reserved = 0;
    int64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;

    bool admissible() const {
        return !queue.empty() && (reserved == 0 || reserved + queue.front()->need <= budget);
    }

    void run(ServeJob &j, hipStream_t s) {
        const qs_job &q = j.spec;
        const int64_t n = j.op->basis->Ns;
        const size_t esz = dtype_size(q.dtype);
        if (q.kind == QS_JOB_EIGSH) {
            const bool herm = j.op->hermitian != 0;
            j.vals.assign(herm ? q.k : 2 * q.k, 0.0);
            if (q.want_vectors) j.out.assign((size_t)n * q.k * esz, 0);
            eigsh_host(*j.op, s, q.dtype, q.k, q.which, q.tol, q.maxiter, q.ncv, j.v0.empty() ? nullptr : j.v0.data(),
                       j.vals.data(), q.want_vectors ? j.out.data() : nullptr, &j.nconv, &j.nmv);
        } else {
            j.out.assign((size_t)n * std::max<int64_t>(1, q.n_times) * esz, 0);
            expm_host(*j.op, s, q.dtype, j.v0.data(), q.n_times, j.scales.data(), q.tol, q.krylov_dim, j.out.data(),
                      &j.nmv);
        }
    }

    void worker(int w) {
        (void)hipSetDevice(0);
        for (;;) {
            ServeJob *j = nullptr;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_work.wait(lk, [&] { return (stop && queue.empty()) || admissible(); });
                if (stop && queue.empty()) return;
                j = queue.front();
                queue.pop_front();
                reserved += j->need;
            }
            j->t_start = now_s();
            try {
                run(*j, streams[w]);
            } catch (const Error &e) {
                j->status = e.code;
                j->err = e.what();
            } catch (const std::bad_alloc &) {
                j->status = QS_EOOM;
                j->err = "host allocation failed";
            } catch (const std::exception &e) {
                j->status = QS_EDEVICE;
                j->err = e.what();
            }
            j->t_end = now_s();
            {
                std::lock_guard<std::mutex> lk(m);
                reserved -= j->need;
                j->done = true;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

// Device bytes a job holds while it runs (Krylov basis, work vectors, outputs).
size_t job_bytes(const qs_job &q, int64_t n) {
    const size_t esz = dtype_size(q.dtype);
    if (q.kind == QS_JOB_EIGSH) {
        const int ncv = q.ncv > 0 ? q.ncv : std::max(2 * q.k + 1, 20);
        return (size_t)n * esz * (size_t)(ncv + 4 + (q.want_vectors ? q.k : 0));
    }
    // basis, output columns, halving midpoints down to the recursion limit
    return (size_t)n * esz * (size_t)(q.krylov_dim + 4 + std::max<int64_t>(1, q.n_times) + 14);
}

}  // namespace qs

struct qs_server {
    qs::Server s;
};

using namespace qs;

namespace {

template <typename F>
int sguard(F &&f) {
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

}  // namespace

extern "C" {

int qs_server_create(int n_workers, int64_t hbm_budget_bytes, qs_server **out) {
    return sguard([&] {
        if (!out) fail(QS_EARG, "out must not be NULL");
        *out = nullptr;
        if (n_workers < 1 || n_workers > 16) fail(QS_EARG, "n_workers must lie in 1:16");
        if (hbm_budget_bytes <= 0) fail(QS_EARG, "the HBM budget must be positive");
        auto h = std::make_unique<qs_server>();
        Server &S = h->s;
        S.budget = (size_t)hbm_budget_bytes;
        for (int w = 0; w < n_workers; w++) {
            hipStream_t st;
            QS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            S.streams.push_back(st);
        }
        for (int w = 0; w < n_workers; w++) S.workers.emplace_back([&S, w] { S.worker(w); });
        *out = h.release();
    });
}

int qs_server_destroy(qs_server *srv) {
    return sguard([&] {
        if (!srv) return;
        Server &S = srv->s;
        {
            std::lock_guard<std::mutex> lk(S.m);
            S.stop = true;
        }
        S.cv_work.notify_all();
        for (auto &t : S.workers) t.join();
        for (auto st : S.streams) (void)hipStreamDestroy(st);
        delete srv;
    });
}

int qs_server_submit(qs_server *srv, const qs_job *job, int64_t *ticket) {
    return sguard([&] {
        if (!srv || !job || !ticket) fail(QS_EARG, "server, job and ticket must not be NULL");
        if (!job->op) fail(QS_EARG, "job.op must not be NULL");
        Server &S = srv->s;
        auto j = std::make_unique<ServeJob>();
        j->spec = *job;
        j->op = &job->op->o;
        const int64_t n = j->op->basis->Ns;
        const size_t esz = dtype_size(job->dtype);
        if (j->op->is_complex && !dtype_complex(job->dtype))
            fail(QS_EARG, "a complex operator needs a complex vector dtype");
        if (job->kind == QS_JOB_EIGSH) {
            if (job->k < 1 || job->k > n) fail(QS_EARG, "k must lie in 1:Ns");
            if (job->v0) j->v0.assign((const char *)job->v0, (const char *)job->v0 + (size_t)n * esz);
        } else if (job->kind == QS_JOB_EXPM) {
            if (!job->v0) fail(QS_EARG, "an expm job needs its initial vector");
            if (!dtype_complex(job->dtype)) fail(QS_EARG, "expm_multiply needs a complex vector dtype");
            if (job->n_times < 0 || (job->n_times > 0 && !job->scales)) fail(QS_EARG, "bad time grid");
            if (job->krylov_dim <= 0) fail(QS_EARG, "krylov_dim must be positive");
            if (job->tol <= 0) fail(QS_EARG, "tol must be positive");
            j->v0.assign((const char *)job->v0, (const char *)job->v0 + (size_t)n * esz);
            j->scales.assign(job->scales, job->scales + 2 * job->n_times);
        } else {
            fail(QS_EARG, "unknown job kind");
        }
        j->spec.v0 = nullptr;
        j->spec.scales = nullptr;
        j->need = job_bytes(*job, n);
        j->t_submit = now_s();
        {
            std::lock_guard<std::mutex> lk(S.m);
            if (S.stop) fail(QS_EARG, "the server is shutting down");
            j->id = S.next_id++;
            *ticket = j->id;
            S.queue.push_back(j.get());
            S.jobs[j->id] = std::move(j);
        }
        S.cv_work.notify_all();
    });
}

int qs_server_wait(qs_server *srv, int64_t ticket, double *vals, void *out, int *nconv, int *n_matvec,
                   double *seconds) {
    return sguard([&] {
        if (!srv) fail(QS_EARG, "server must not be NULL");
        Server &S = srv->s;
        std::unique_ptr<ServeJob> j;
        {
            std::unique_lock<std::mutex> lk(S.m);
            auto it = S.jobs.find(ticket);
            if (it == S.jobs.end()) fail(QS_EARG, "unknown ticket");
            ServeJob *p = it->second.get();
            S.cv_done.wait(lk, [&] { return p->done; });
            j = std::move(it->second);
            S.jobs.erase(it);
        }
        if (seconds) {
            seconds[0] = j->t_start - j->t_submit;
            seconds[1] = j->t_end - j->t_start;
            seconds[2] = j->t_end - j->t_submit;
        }
        if (n_matvec) *n_matvec = j->nmv;
        if (nconv) *nconv = j->nconv;
        if (vals && !j->vals.empty()) memcpy(vals, j->vals.data(), j->vals.size() * sizeof(double));
        if (out && !j->out.empty()) memcpy(out, j->out.data(), j->out.size());
        if (j->status != QS_OK) fail(j->status, j->err);
    });
}

}  // extern "C"
