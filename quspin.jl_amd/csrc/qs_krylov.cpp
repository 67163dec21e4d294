// qs_krylov.cpp -- device-resident Krylov solvers driven from the host.
//
// eigsh_device: replaces Arpack.eigs behind eigsh (operators.jl:3367-3408).
//   Thick-restart Lanczos (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000)
//   602) with two-pass classical Gram-Schmidt against the whole resident
//   basis, FP64 accumulation, and ARPACK's symmetric convergence test
//   |beta_m y_m,i| <= tol * max(eps^(2/3), |theta_i|).  The Krylov basis stays
//   in HBM for the whole solve; only the small projected matrix visits the host.
// expm_times_device: _krylov_expmv_times (operators.jl:1805-1893) with
//   _krylov_factorization (:1655-1731), _krylov_project (:1733-1756) and the
//   halving recursion of _krylov_expmv_vector (:1758-1803).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "qs_internal.h"

namespace qs {

namespace {

const double kEps = 2.220446049250313e-16;

struct Krylov {
    int dtype;
    int64_t n;
    size_t esz;
    SBuf<char> V;   // n x cols
    int cols = 0;
    Krylov(int dt, int64_t n_, int c, hipStream_t s) : dtype(dt), n(n_), esz(dtype_size(dt)), cols(c) {
        V.alloc((size_t)n * c * esz, s);
    }
    void *col(int j) const { return (void *)(V.get() + (size_t)j * n * esz); }
    void swap(Krylov &o) {
        std::swap(V.p, o.V.p);
        std::swap(V.n, o.V.n);
        std::swap(cols, o.cols);
    }
};

// Second basis of a thick-restart solve: the restart's recombination writes
// into it and the two swap, instead of recombining into scratch and copying
// the kept columns back.  Allocated at the first restart when free device
// memory holds three times its size; otherwise the restart recombines
// through scratch.
struct SpareBasis {
    std::unique_ptr<Krylov> K;
    bool failed = false;
    Krylov *get(int dtype, int64_t n, int cols, hipStream_t s) {
        if (!K && !failed) {
            // only where it leaves room for the rest of the solve
            size_t fr = 0, tot = 0;
            const size_t bytes = (size_t)n * cols * dtype_size(dtype);
            if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < 3 * bytes) {
                (void)hipGetLastError();
                failed = true;
                return nullptr;
            }
            try {
                K = std::make_unique<Krylov>(dtype, n, cols, s);
            } catch (const Error &) {
                failed = true;
            }
        }
        return K.get();
    }
};

void matvec(OpHost &op, int dtype, const void *x, void *y) {
    launch_matvec(op, dtype, x, y, 1.0, 0.0, 0.0, 0.0, false, false, false, op_stream(op));
}

// w (already holding A v_j) orthogonalised against V[:, 0..j] twice;
// returns h (sum of both passes) and ||w||.
double orthogonalize(Krylov &K, int j, void *w, VecWork &wk, std::vector<cd> &h, hipStream_t s) {
    std::vector<cd> h2;
    gram_dots(K.dtype, K.V.get(), K.n, j + 1, w, wk, s, h);
    gram_update(K.dtype, K.V.get(), K.n, j + 1, h, w, wk, s);
    gram_dots(K.dtype, K.V.get(), K.n, j + 1, w, wk, s, h2);
    double nrm2 = gram_update(K.dtype, K.V.get(), K.n, j + 1, h2, w, wk, s);
    for (int i = 0; i <= j; i++) h[i] += h2[i];
    return std::sqrt(nrm2);
}

bool better(int which, double a, double b) {
    // true if Ritz value a is more wanted than b
    switch (which) {
    case QS_WHICH_SA: return a < b;
    case QS_WHICH_LA: return a > b;
    case QS_WHICH_SM: return std::fabs(a) < std::fabs(b);
    case QS_WHICH_LM: return std::fabs(a) > std::fabs(b);
    }
    return a < b;
}

// indices of the `want` most wanted Ritz values (BE: alternate both ends,
// one more from the high end when want is odd)
std::vector<int> select(const std::vector<double> &theta, int which, int want) {
    int m = (int)theta.size();
    std::vector<int> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    if (which == QS_WHICH_BE) {
        std::sort(idx.begin(), idx.end(), [&](int x, int y) { return theta[x] < theta[y]; });
        std::vector<int> out;
        int lo = 0, hi = m - 1;
        bool take_hi = (want % 2 == 1);
        while ((int)out.size() < want && lo <= hi) {
            if (take_hi) out.push_back(idx[hi--]);
            else out.push_back(idx[lo++]);
            take_hi = !take_hi;
        }
        return out;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return better(which, theta[x], theta[y]); });
    idx.resize(std::min(want, m));
    return idx;
}

// FP32-stored Ritz vectors Y (n x k): Rayleigh-Ritz in FP64 on span(Y).
// Y is widened to FP64, W = A Y is formed by FP64 products, and the block
// Gram products G = Y^H W and M = Y^H Y run on the matrix cores; the k x k
// problem G c = lambda M c (Cholesky of M) gives the reported eigenvalues,
// accurate to second order in the FP32 residual, and its vectors recombine Y
// into the returned FP32 Ritz vectors.
void rayleigh_ritz_fp64(OpHost &op, int dtype, int k, int which, const void *Yf, int64_t n, VecWork &wk,
                        EigshResult &res, void *vecs_dev) {
    hipStream_t s = op_stream(op);
    const int dt64 = dtype == QS_F32 ? QS_F64 : QS_C128;
    const size_t e64 = dtype_size(dt64);
    SBuf<char> Y64((size_t)n * k * e64, s), W64((size_t)n * k * e64, s);
    vec_convert(dtype, Yf, dt64, Y64.get(), n * k, s);
    for (int c = 0; c < k; c++) {
        launch_matvec(op, dt64, Y64.get() + (size_t)c * n * e64, W64.get() + (size_t)c * n * e64, 1.0, 0.0, 0.0,
                      0.0, false, false, false, s);
        res.nmatvec++;
    }
    std::vector<cd> G, M;
    gram_block(dt64, Y64.get(), n, k, W64.get(), k, wk, s, G);
    gram_block(dt64, Y64.get(), n, k, Y64.get(), k, wk, s, M);
    auto g = [&](std::vector<cd> &X, int i, int j) -> cd & { return X[(size_t)i + (size_t)k * j]; };
    // M = L L^H
    std::vector<cd> Lm((size_t)k * k, cd(0, 0));
    auto l = [&](int i, int j) -> cd & { return Lm[(size_t)i * k + j]; };
    for (int j = 0; j < k; j++) {
        double dsum = g(M, j, j).real();
        for (int q = 0; q < j; q++) dsum -= std::norm(l(j, q));
        if (!(dsum > 0)) fail(QS_ENOCONV, "eigsh: FP32 Ritz vectors are linearly dependent");
        l(j, j) = std::sqrt(dsum);
        for (int i = j + 1; i < k; i++) {
            cd acc = g(M, i, j);
            for (int q = 0; q < j; q++) acc -= l(i, q) * std::conj(l(j, q));
            l(i, j) = acc / l(j, j);
        }
    }
    // A = L^-1 G L^-H (Hermitian part)
    std::vector<cd> X((size_t)k * k), A((size_t)k * k);
    for (int c = 0; c < k; c++)          // X = L^-1 G, column by column
        for (int i = 0; i < k; i++) {
            cd acc = g(G, i, c);
            for (int q = 0; q < i; q++) acc -= l(i, q) * X[(size_t)q * k + c];
            X[(size_t)i * k + c] = acc / l(i, i);
        }
    for (int r = 0; r < k; r++)          // A = X L^-H: solve A L^H = X row by row
        for (int j = 0; j < k; j++) {
            cd acc = X[(size_t)r * k + j];
            for (int q = 0; q < j; q++) acc -= A[(size_t)r * k + q] * std::conj(l(j, q));
            A[(size_t)r * k + j] = acc / std::conj(l(j, j));
        }
    for (int i = 0; i < k; i++)
        for (int j = i; j < k; j++) {
            const cd h = 0.5 * (A[(size_t)i * k + j] + std::conj(A[(size_t)j * k + i]));
            A[(size_t)i * k + j] = h;
            A[(size_t)j * k + i] = std::conj(h);
        }
    std::vector<double> lam;
    std::vector<cd> U;
    herm_eig(k, A, lam, U);
    std::vector<int> ord(k);
    std::iota(ord.begin(), ord.end(), 0);
    auto key = [&](int i) {
        switch (which) {
        case QS_WHICH_LA: return -lam[i];
        case QS_WHICH_SM: return std::fabs(lam[i]);
        case QS_WHICH_LM: return -std::fabs(lam[i]);
        default: return lam[i];
        }
    };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key(a) < key(b); });
    for (int c = 0; c < k; c++) res.vals[c] = lam[ord[c]];
    if (vecs_dev) {
        // coefficients L^-H U of the refined vectors, in the reported order
        std::vector<cd> Cr((size_t)k * k);
        for (int c = 0; c < k; c++)
            for (int i = k - 1; i >= 0; i--) {
                cd acc = U[(size_t)i * k + ord[c]];
                for (int q = i + 1; q < k; q++) acc -= std::conj(l(q, i)) * Cr[(size_t)c * k + q];
                Cr[(size_t)c * k + i] = acc / std::conj(l(i, i));
            }
        basis_combine(dtype, Yf, n, k, Cr, k, vecs_dev, s);
    }
}

}  // namespace

void eigsh_device(OpHost &op, int dtype, int k, int which, double tol, int maxiter, int ncv, const void *v0_dev,
                  EigshResult &res, void *vecs_dev) {
    const int64_t n = op.basis->Ns;
    hipStream_t s = op_stream(op);
    if (!op.hermitian) fail(QS_EARG, "eigsh on the device needs a Hermitian operator");
    if (tol <= 0) tol = kEps;
    if (maxiter <= 0) maxiter = 300;
    if (ncv <= 0) ncv = std::max(20, 2 * k + 1);
    ncv = (int)std::min<int64_t>(ncv, n);
    if (ncv <= k) fail(QS_EARG, "ncv must exceed k");
    const double eps23 = std::pow(kEps, 2.0 / 3.0);
    Krylov K(dtype, n, ncv + 1, s);
    SpareBasis spare;
    VecWork wk(s);
    // starting vector
    if (v0_dev) {
        QS_HIP(hipMemcpyAsync(K.col(0), v0_dev, (size_t)n * K.esz, hipMemcpyDeviceToDevice, s));
    } else {
        vec_random(dtype, K.col(0), n, 0x51a7ULL, s);
    }
    double nrm = std::sqrt(vec_norm2(dtype, K.col(0), n, wk, s));
    if (!(nrm > 0)) fail(QS_EARG, "initial vector must be nonzero");
    vec_scale_copy(dtype, K.col(0), K.col(0), n, cd(1.0 / nrm, 0), s);

    const int m = ncv;
    std::vector<double> T((size_t)(m + 1) * (m + 1), 0.0);
    auto t = [&](int i, int j) -> double & { return T[(size_t)i * (m + 1) + j]; };
    int kept = 0;
    double beta_last = 0;
    std::vector<double> theta, Y;
    std::vector<int> wanted;
    int nconv = 0;
    uint64_t reseed = 1;
    int kw = k;                  // Ritz values converged: k, then k + 1 while verifying
    bool have_record = false;
    int verifications = 0;
    std::vector<double> record;  // the k wanted values of the last converged round
    // per-step scalars of one Lanczos cycle, device side: two Gram passes
    // (m + 1 complex each), |w|^2 after both and |w'|^2 after the first
    const size_t per = (size_t)4 * (m + 1) + 4;
    SBuf<double> d(per * m, s);
    std::vector<double> host(per * m);
    vec_reserve(wk, n, m + 1);
    for (int iter = 0; iter < maxiter; iter++) {
        int j0 = kept;
        while (j0 < m) {
            // queue steps j0 .. m-1 with no host round trip (a small
            // assembled sector: all of them in one single-workgroup launch)
            ensure_assembled(op, s);
            if (krylov_small_ok(op, m + 1)) {
                krylov_small(op, dtype, SMALL_DGKS, m, j0, K.V.get(), nullptr, d.get(), (int)per, s);
            } else {
                for (int j = j0; j < m; j++) {
                    void *w = K.col(j + 1);
                    double *h1 = d.get() + per * j, *h2 = h1 + 2 * (m + 1), *n2 = h2 + 2 * (m + 1), *n2a = n2 + 2;
                    matvec(op, dtype, K.col(j), w);
                    // ARPACK's DGKS step: the second pass runs only if the first
                    // left |w'| <= 0.717 |w| (decided on the device)
                    gram_dgks_dev(dtype, K.V.get(), n, j + 1, w, wk, h1, h2, n2a, n2, s);
                    vec_axpby_dev(dtype, n2, CO_INV_SQRT, w, cd(0, 0), w, n, s);
                }
            }
            QS_HIP(hipMemcpyAsync(host.data() + per * j0, d.get() + per * j0, sizeof(double) * per * (m - j0),
                                  hipMemcpyDeviceToHost, s));
            QS_HIP(hipStreamSynchronize(s));
            // the step-by-step bookkeeping, in order; a breakdown discards the
            // steps queued after it
            int next = m;
            for (int j = j0; j < m; j++) {
                res.nmatvec++;
                const double *h1 = host.data() + per * j, *h2 = h1 + 2 * (m + 1);
                std::vector<cd> h(j + 1);
                for (int i = 0; i <= j; i++) h[i] = cd(h1[2 * i], h1[2 * i + 1]) + cd(h2[2 * i], h2[2 * i + 1]);
                double beta = std::sqrt(h2[2 * (m + 1)]);
                for (int i = kept; i <= j; i++) {
                    // the Rayleigh quotient entries of the new column; entries
                    // below `kept` are the arrow couplings set at restart
                    t(i, j) = h[i].real();
                    t(j, i) = h[i].real();
                }
                double tnorm = 0;
                for (double v : T) tnorm = std::max(tnorm, std::fabs(v));
                const bool brk = beta <= 1e3 * kEps * std::max(1.0, tnorm);
                if (brk) {
                    // invariant subspace: continue with a fresh direction
                    beta = 0;
                    void *w = K.col(j + 1);
                    vec_random(dtype, w, n, 0x9e37ULL + (reseed++), s);
                    std::vector<cd> hh;
                    double b2 = orthogonalize(K, j, w, wk, hh, s);
                    vec_scale_copy(dtype, w, w, n, cd(1.0 / b2, 0), s);
                }
                if (j + 1 < m) {
                    t(j, j + 1) = beta;
                    t(j + 1, j) = beta;
                }
                beta_last = beta;
                if (brk) {
                    next = j + 1;
                    break;
                }
            }
            j0 = next;
        }
        // Rayleigh-Ritz on the m x m projected matrix
        std::vector<double> A((size_t)m * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) A[(size_t)i * m + j] = t(i, j);
        sym_eig(m, A, theta, Y);
        std::vector<int> wanted_all = select(theta, which, kw);
        // the Ritz value beyond the k wanted (verification rounds) only has to
        // settle enough to tell whether it belongs among them
        const double tol_extra = std::max(tol, 1e-6);
        int nconv_all = 0;
        for (size_t i = 0; i < wanted_all.size(); i++) {
            const int idx = wanted_all[i];
            double r = std::fabs(beta_last * Y[(size_t)(m - 1) * m + idx]);
            if (r <= ((int)i < k ? tol : tol_extra) * std::max(eps23, std::fabs(theta[idx]))) nconv_all++;
        }
        wanted.assign(wanted_all.begin(), wanted_all.begin() + std::min<size_t>(k, wanted_all.size()));
        nconv = 0;
        for (int idx : wanted) {
            double r = std::fabs(beta_last * Y[(size_t)(m - 1) * m + idx]);
            if (r <= tol * std::max(eps23, std::fabs(theta[idx]))) nconv++;
        }
        if (nconv_all >= kw) {
            // Converged.  A single-vector Krylov space holds one vector of each
            // eigenspace, so a repeated eigenvalue shows only through rounding
            // and can be missed.  So, once: restart with the converged Ritz
            // vectors beside a random direction orthogonal to them (its
            // Krylov block starts fresh, uncoupled) and converge one Ritz value
            // more than asked.  The k wanted values are accepted when a
            // verification round leaves them unchanged as a multiset.
            std::vector<double> now;
            for (int idx : wanted) now.push_back(theta[idx]);
            std::sort(now.begin(), now.end());
            if (have_record) {
                bool same = now.size() == record.size();
                const double scale = std::max(1.0, std::fabs(now.empty() ? 0.0 : now.back()));
                for (size_t i = 0; same && i < now.size(); i++)
                    same = std::fabs(now[i] - record[i]) <= std::max(1e3 * kEps, tol) * scale;
                if (same) break;
            }
            // (with k = 1 a missed copy of the lowest value cannot change the
            // answer, so only k >= 2 verifies)
            const bool can = k >= 2 && which != QS_WHICH_BE && m - (k + 1) >= 2 && verifications < k + 2;
            if (!can || iter == maxiter - 1) break;
            verifications++;
            record = now;
            have_record = true;
            const int kc = (int)wanted_all.size();
            std::vector<cd> C((size_t)m * kc);
            for (int c = 0; c < kc; c++)
                for (int j = 0; j < m; j++) C[(size_t)c * m + j] = cd(Y[(size_t)j * m + wanted_all[c]], 0);
            if (Krylov *K2 = spare.get(dtype, n, ncv + 1, s)) {
                basis_combine(dtype, K.V.get(), n, m, C, kc, K2->V.get(), s);
                K.swap(*K2);
            } else {
                Krylov tmp(dtype, n, kc, s);
                basis_combine(dtype, K.V.get(), n, m, C, kc, tmp.V.get(), s);
                QS_HIP(hipMemcpyAsync(K.col(0), tmp.V.get(), (size_t)n * kc * K.esz, hipMemcpyDeviceToDevice, s));
            }
            void *w = K.col(kc);
            vec_random(dtype, w, n, 0x7e51ULL + (reseed++), s);
            std::vector<cd> hh;
            double b2 = orthogonalize(K, kc - 1, w, wk, hh, s);
            vec_scale_copy(dtype, w, w, n, cd(1.0 / b2, 0), s);
            std::fill(T.begin(), T.end(), 0.0);
            for (int c = 0; c < kc; c++) t(c, c) = theta[wanted_all[c]];
            kept = kc;
            kw = k + 1;
            continue;
        }
        if (iter == maxiter - 1) break;
        // thick restart: keep the wanted Ritz vectors plus as many runners-up
        int keep = std::min(m - 1, kw + (m - kw) / 2);
        std::vector<int> sel = select(theta, which, keep);
        std::vector<cd> C((size_t)m * keep);
        for (int c = 0; c < keep; c++)
            for (int j = 0; j < m; j++) C[(size_t)c * m + j] = cd(Y[(size_t)j * m + sel[c]], 0);
        if (Krylov *K2 = spare.get(dtype, n, ncv + 1, s)) {
            basis_combine(dtype, K.V.get(), n, m, C, keep, K2->V.get(), s);
            QS_HIP(hipMemcpyAsync(K2->col(keep), K.col(m), (size_t)n * K.esz, hipMemcpyDeviceToDevice, s));
            K.swap(*K2);
        } else {
            Krylov tmp(dtype, n, keep, s);
            basis_combine(dtype, K.V.get(), n, m, C, keep, tmp.V.get(), s);
            QS_HIP(hipMemcpyAsync(K.col(0), tmp.V.get(), (size_t)n * keep * K.esz, hipMemcpyDeviceToDevice, s));
            QS_HIP(hipMemcpyAsync(K.col(keep), K.col(m), (size_t)n * K.esz, hipMemcpyDeviceToDevice, s));
        }
        std::fill(T.begin(), T.end(), 0.0);
        for (int c = 0; c < keep; c++) {
            t(c, c) = theta[sel[c]];
            double sc = beta_last * Y[(size_t)(m - 1) * m + sel[c]];
            t(c, keep) = sc;
            t(keep, c) = sc;
        }
        kept = keep;
    }
    // order the output the way operators.jl:3393-3404 does
    std::vector<int> out = wanted;
    if (which == QS_WHICH_SA) std::sort(out.begin(), out.end(), [&](int a, int b) { return theta[a] < theta[b]; });
    else if (which == QS_WHICH_LA) std::sort(out.begin(), out.end(), [&](int a, int b) { return theta[a] > theta[b]; });
    else if (which == QS_WHICH_SM)
        std::stable_sort(out.begin(), out.end(), [&](int a, int b) { return std::fabs(theta[a]) < std::fabs(theta[b]); });
    else if (which == QS_WHICH_LM)
        std::stable_sort(out.begin(), out.end(), [&](int a, int b) { return std::fabs(theta[a]) > std::fabs(theta[b]); });
    else std::sort(out.begin(), out.end(), [&](int a, int b) { return theta[a] < theta[b]; });
    res.vals.resize(k);
    for (int i = 0; i < k; i++) res.vals[i] = theta[out[i]];
    res.nconv = nconv;
    const bool fp32 = dtype == QS_F32 || dtype == QS_C64;
    if (vecs_dev || fp32) {
        std::vector<cd> C((size_t)m * k);
        for (int c = 0; c < k; c++)
            for (int j = 0; j < m; j++) C[(size_t)c * m + j] = cd(Y[(size_t)j * m + out[c]], 0);
        Krylov tmp(dtype, n, (vecs_dev && !fp32) ? 0 : k, s);
        void *dst = (vecs_dev && !fp32) ? vecs_dev : tmp.V.get();
        basis_combine(dtype, K.V.get(), n, m, C, k, dst, s);
        if (fp32) rayleigh_ritz_fp64(op, dtype, k, which, dst, n, wk, res, vecs_dev);
    }
    QS_HIP(hipStreamSynchronize(s));
}

// -------------------------------------------------------------- expm ------
namespace {

}  // namespace

// The reference's Hermitian path is Lanczos without reorthogonalisation
// (operators.jl:1686-1700).  With FP32-stored vectors orthogonality is lost
// within a few steps and long-time results drift by O(1e-1); the FP32 form
// therefore always takes the two-pass Gram-Schmidt (Arnoldi) path, whose
// reorthogonalisation runs in FP64.
bool lanczos_ok(const OpHost &op, int dtype) { return op.hermitian != 0 && (dtype == QS_F64 || dtype == QS_C128); }

namespace {

struct Factorization {
    std::unique_ptr<Krylov> K;   // own basis, or null when a FactPlan holds it
    const void *basis = nullptr;
    KrylovProj P;
};

// A factorization whose buffers stay fixed across the calls of one halving
// recursion, so its launch sequence (every step, scalars on the device) is
// captured once in a HIP graph and replayed by each later call; the input is
// copied into `vin` first.  Falls back to direct launches if capture fails.
struct FactPlan {
    int64_t n = 0;
    int steps = 0;
    SBuf<char> vin, V;
    SBuf<double> d;
    hipGraphExec_t exec = nullptr;
    bool no_graph = false;
    int uses = 0;
    ~FactPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

// every step of the factorization of vin into V (n x (steps + 1)), scalars in d
void queue_factorization(OpHost &op, int dtype, int64_t n, int steps, bool hermitian, const void *vin, char *V,
                         double *dd, VecWork &wk, hipStream_t s) {
    const size_t esz = dtype_size(dtype);
    auto col = [&](int j) { return (void *)(V + (size_t)j * n * esz); };
    FactScalars L(steps);
    ensure_assembled(op, s);
    if (krylov_small_ok(op, steps + 1)) {
        // a small assembled sector: the whole factorization in one launch
        krylov_small(op, dtype, hermitian ? SMALL_LANCZOS : SMALL_ARNOLDI, steps, 0, V, vin, dd, 0, s);
        return;
    }
    vec_norm2_dev(dtype, vin, n, wk, dd + L.beta2(), s);
    vec_axpby_dev(dtype, dd + L.beta2(), CO_INV_SQRT, vin, cd(0, 0), col(0), n, s);
    for (int c = 0; c < steps; c++) {
        void *w = col(c + 1);
        double *h1 = dd + L.h1(c), *h2 = dd + L.h2(c), *n2 = dd + L.nrm2(c);
        const int nbs = (int)((n + 255) / 256);
        if (hermitian && (size_t)2 * nbs <= wk.partial.n &&
            matvec_lanczos(op, dtype, col(c), w, c > 0 ? col(c - 1) : nullptr, c > 0 ? dd + L.nrm2(c - 1) : nullptr,
                           CO_NEG_SQRT, wk.partial.get(), s)) {
            // assembled: the product also subtracts beta v_{c-1} and leaves
            // the per-block partials of <v_c, w>
            vec_reduce_dev(wk.partial.get(), nbs, h1, s);
            gram_update_dev(dtype, col(c), n, 1, h1, w, wk, n2, s);
        } else if (hermitian) {
            // three-term Lanczos without reorthogonalisation
            matvec(op, dtype, col(c), w);
            if (c > 0) vec_axpby_dev(dtype, dd + L.nrm2(c - 1), CO_NEG_SQRT, col(c - 1), cd(1, 0), w, n, s);
            gram_dots_dev(dtype, col(c), n, 1, w, wk, h1, s);
            gram_update_dev(dtype, col(c), n, 1, h1, w, wk, n2, s);
        } else {
            matvec(op, dtype, col(c), w);
            // Arnoldi, two Gram-Schmidt passes
            gram_dots_dev(dtype, V, n, c + 1, w, wk, h1, s);
            gram_update_dev(dtype, V, n, c + 1, h1, w, wk, nullptr, s);
            gram_dots_dev(dtype, V, n, c + 1, w, wk, h2, s);
            gram_update_dev(dtype, V, n, c + 1, h2, w, wk, n2, s);
        }
        if (c + 1 < steps) vec_axpby_dev(dtype, n2, CO_INV_SQRT, w, cd(0, 0), w, n, s);
    }
}

// operators.jl:1655-1731.  Every step is queued on the stream with its
// scalars (h, |w|^2) kept in device memory; one copy brings them to the host
// at the end, where the reference's per-step bookkeeping and breakdown test
// run in the original order (KrylovProj::replay).  The test at step c only
// reads quantities of steps <= c, so testing afterwards stops at the same
// step; columns queued past a breakdown are never used.
void factorize(OpHost &op, int dtype, const void *v, int krylov_dim, Factorization &F, VecWork &wk, int &nmv) {
    hipStream_t s = op_stream(op);
    const int64_t n = op.basis->Ns;
    KrylovProj &P = F.P;
    P.dim = n;
    P.hermitian = lanczos_ok(op, dtype);
    P.steps = (int)std::min<int64_t>(krylov_dim, n);
    F.K = std::make_unique<Krylov>(dtype, n, P.steps + 1, s);
    F.basis = F.K->V.get();
    P.H.assign((size_t)(P.steps + 1) * P.steps, cd(0, 0));
    P.completed = P.steps;
    P.terminal = 0;
    const int steps = P.steps;
    vec_reserve(wk, n, steps + 1);
    FactScalars L(steps);
    SBuf<double> d(L.total(), s);
    queue_factorization(op, dtype, n, steps, P.hermitian, v, F.K->V.get(), d.get(), wk, s);
    std::vector<double> host(L.total());
    QS_HIP(hipMemcpyAsync(host.data(), d.get(), sizeof(double) * host.size(), hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    nmv += P.replay(L, host);
}

// The same through a FactPlan: the second and later calls replay one graph.
void factorize_plan(OpHost &op, int dtype, const void *v, int krylov_dim, FactPlan &plan, Factorization &F,
                    VecWork &wk, int &nmv) {
    hipStream_t s = op_stream(op);
    const int64_t n = op.basis->Ns;
    const size_t esz = dtype_size(dtype);
    KrylovProj &P = F.P;
    P.dim = n;
    P.hermitian = lanczos_ok(op, dtype);
    P.steps = (int)std::min<int64_t>(krylov_dim, n);
    P.H.assign((size_t)(P.steps + 1) * P.steps, cd(0, 0));
    P.completed = P.steps;
    P.terminal = 0;
    const int steps = P.steps;
    FactScalars L(steps);
    if (!plan.V.get()) {
        plan.n = n;
        plan.steps = steps;
        plan.vin.alloc((size_t)n * esz, s);
        plan.V.alloc((size_t)n * (steps + 1) * esz, s);
        plan.d.alloc(L.total(), s);
        vec_reserve(wk, n, steps + 1);
    }
    F.basis = plan.V.get();
    QS_HIP(hipMemcpyAsync(plan.vin.get(), v, (size_t)n * esz, hipMemcpyDeviceToDevice, s));
    auto queue = [&] {
        queue_factorization(op, dtype, n, steps, P.hermitian, plan.vin.get(), plan.V.get(), plan.d.get(), wk, s);
    };
    if (!plan.exec && !plan.no_graph && plan.uses >= 1) {
        // only when the operator opted in (qs_op_set_graphs), and only on a
        // non-blocking stream: while a capture is open the runtime refuses
        // legacy-default-stream work from every thread
        unsigned flags = 0;
        if (!op.graphs || hipStreamGetFlags(s, &flags) != hipSuccess || !(flags & hipStreamNonBlocking)) {
            (void)hipGetLastError();
            plan.no_graph = true;
        }
    }
    if (!plan.exec && !plan.no_graph && plan.uses >= 1) {
        // the first call ran directly (lazy tables, workspaces and assembly are
        // in place); capture this call's launches
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            try {
                queue();
            } catch (const Error &) {
                ok = false;
            }
            ok = (hipStreamEndCapture(s, &g) == hipSuccess) && ok && g;
            ok = ok && hipGraphInstantiate(&plan.exec, g, nullptr, nullptr, 0) == hipSuccess;
            if (g) (void)hipGraphDestroy(g);
        }
        if (!ok) {
            (void)hipGetLastError();
            plan.exec = nullptr;
            plan.no_graph = true;
        }
    }
    if (plan.exec) QS_HIP(hipGraphLaunch(plan.exec, s));
    else queue();
    plan.uses++;
    std::vector<double> host(L.total());
    QS_HIP(hipMemcpyAsync(host.data(), plan.d.get(), sizeof(double) * host.size(), hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    nmv += P.replay(L, host);
}

// operators.jl:1733-1756: result = beta * V[:, :c] exp(scale H_c)[:, 0]
double project(OpHost &op, int dtype, Factorization &F, cd scale, void *out, VecWork &wk) {
    std::vector<cd> coef;
    const double est = F.P.expm_column(scale, coef);
    basis_combine(dtype, F.basis, F.P.dim, F.P.completed, coef, 1, out, op_stream(op));
    return est;
}

// operators.jl:1758-1803
void expmv_vector(OpHost &op, int dtype, const void *v, cd scale, double tol, int krylov_dim, int depth, void *out,
                  VecWork &wk, FactPlan &plan, int &nmv) {
    hipStream_t s = op_stream(op);
    const int64_t n = op.basis->Ns;
    if (tol <= 0) fail(QS_EARG, "tol must be positive");
    if (scale == cd(0, 0) || vec_norm2(dtype, v, n, wk, s) == 0) {
        QS_HIP(hipMemcpyAsync(out, v, (size_t)n * dtype_size(dtype), hipMemcpyDeviceToDevice, s));
        return;
    }
    // the plan's basis is reused by the nested calls: this call is done with
    // it once `out` holds the projection
    Factorization F;
    factorize_plan(op, dtype, v, krylov_dim, plan, F, wk, nmv);
    double est = project(op, dtype, F, scale, out, wk);
    double nr = std::sqrt(vec_norm2(dtype, out, n, wk, s));
    if (est > tol * std::max(1.0, nr)) {
        if (depth >= 14) fail(QS_ENOCONV, "Krylov exponential action did not converge; increase krylov_dim");
        SBuf<char> mid((size_t)n * dtype_size(dtype), s);
        expmv_vector(op, dtype, v, scale / 2.0, tol / 2, krylov_dim, depth + 1, mid.get(), wk, plan, nmv);
        expmv_vector(op, dtype, mid.get(), scale / 2.0, tol / 2, krylov_dim, depth + 1, out, wk, plan, nmv);
    }
}

}  // namespace

void expm_times_device(OpHost &op, int dtype, const void *v, const std::vector<cd> &scales, double tol,
                       int krylov_dim, void *out, int &nmatvec) {
    hipStream_t s = op_stream(op);
    const int64_t n = op.basis->Ns;
    const size_t esz = dtype_size(dtype);
    if (!dtype_complex(dtype)) fail(QS_EARG, "expm_multiply needs a complex vector dtype");
    if (tol <= 0) fail(QS_EARG, "tol must be positive");
    if (krylov_dim <= 0) fail(QS_EARG, "krylov_dim must be positive");
    VecWork wk(s);
    auto outcol = [&](size_t j) { return (void *)((char *)out + j * n * esz); };
    bool all_zero = true;
    for (auto &z : scales) all_zero &= z == cd(0, 0);
    if (scales.empty()) return;
    if (all_zero || vec_norm2(dtype, v, n, wk, s) == 0) {
        for (size_t j = 0; j < scales.size(); j++)
            QS_HIP(hipMemcpyAsync(outcol(j), v, n * esz, hipMemcpyDeviceToDevice, s));
        QS_HIP(hipStreamSynchronize(s));
        return;
    }
    Factorization F;
    factorize(op, dtype, v, krylov_dim, F, wk, nmatvec);
    FactPlan plan;   // the halving recursions' factorizations (graph-replayed)
    const int c = F.P.completed;
    if (F.P.hermitian) {
        std::vector<cd> coefs;
        std::vector<double> ests;
        F.P.lanczos_times(scales, coefs, ests);
        basis_combine(dtype, F.K->V.get(), n, c, coefs, (int)scales.size(), out, s);
        for (size_t j = 0; j < scales.size(); j++) {
            if (scales[j] == cd(0, 0)) {
                QS_HIP(hipMemcpyAsync(outcol(j), v, n * esz, hipMemcpyDeviceToDevice, s));
                continue;
            }
            double nr = std::sqrt(vec_norm2(dtype, outcol(j), n, wk, s));
            if (ests[j] > tol * std::max(1.0, nr))
                expmv_vector(op, dtype, v, scales[j], tol, krylov_dim, 0, outcol(j), wk, plan, nmatvec);
        }
    } else {
        for (size_t j = 0; j < scales.size(); j++) {
            if (scales[j] == cd(0, 0)) {
                QS_HIP(hipMemcpyAsync(outcol(j), v, n * esz, hipMemcpyDeviceToDevice, s));
                continue;
            }
            double est = project(op, dtype, F, scales[j], outcol(j), wk);
            double nr = std::sqrt(vec_norm2(dtype, outcol(j), n, wk, s));
            if (est > tol * std::max(1.0, nr))
                expmv_vector(op, dtype, v, scales[j], tol, krylov_dim, 0, outcol(j), wk, plan, nmatvec);
        }
    }
    QS_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------- small projected problem --
void KrylovProj::lanczos_entry(int col, cd alpha, double tn) {
    H[(size_t)col * steps + col] = alpha;
    H[(size_t)(col + 1) * steps + col] = tn;
    if (col + 1 < steps) H[(size_t)col * steps + col + 1] = tn;
}

void KrylovProj::arnoldi_column(int col, const std::vector<cd> &h1, const std::vector<cd> &h2, double tn) {
    for (int r = 0; r <= col; r++) H[(size_t)r * steps + col] = h1[r] + h2[r];
    H[(size_t)(col + 1) * steps + col] = tn;
}

bool KrylovProj::breakdown(int col, double tn, double f) {
    terminal = tn;
    if (tn <= 100 * kEps * std::max(1.0, std::sqrt(f))) {
        completed = col + 1;
        return true;
    }
    return false;
}

int KrylovProj::replay(const FactScalars &L, const std::vector<double> &d) {
    beta = std::sqrt(d[L.beta2()]);
    double f = 0;   // squared Frobenius norm of the entries set so far
    for (int col = 0; col < steps; col++) {
        const double tn = std::sqrt(d[L.nrm2(col)]);
        auto hv = [&](size_t off, int count) {
            std::vector<cd> h(count);
            for (int i = 0; i < count; i++) h[i] = cd(d[off + 2 * i], d[off + 2 * i + 1]);
            return h;
        };
        if (hermitian) {
            const cd alpha(d[L.h1(col)], d[L.h1(col) + 1]);
            lanczos_entry(col, alpha, tn);
            f += std::norm(alpha) + tn * tn + (col + 1 < steps ? tn * tn : 0.0);
        } else {
            arnoldi_column(col, hv(L.h1(col), col + 1), hv(L.h2(col), col + 1), tn);
            for (int r = 0; r <= col; r++) f += std::norm(h(r, col));
            f += tn * tn;
        }
        if (breakdown(col, tn, f)) return col + 1;
    }
    return steps;
}

double KrylovProj::expm_column(cd scale, std::vector<cd> &coef) const {
    const int c = completed;
    if (hermitian && scale != cd(0, 0)) {
        // real symmetric tridiagonal projection: through its eigenvectors
        std::vector<double> est;
        lanczos_times({scale}, coef, est);
        return est[0];
    }
    std::vector<cd> A((size_t)c * c), E;
    for (int i = 0; i < c; i++)
        for (int j = 0; j < c; j++) A[(size_t)i * c + j] = scale * h(i, j);
    expm_dense(c, A, E);
    coef.assign(c, cd(0, 0));
    for (int i = 0; i < c; i++) coef[i] = beta * E[(size_t)i * c + 0];
    const bool full = completed == steps && completed < dim;
    return full ? std::abs(scale) * terminal * beta * std::abs(E[(size_t)(c - 1) * c + 0]) : 0.0;
}

void KrylovProj::lanczos_times(const std::vector<cd> &scales, std::vector<cd> &coefs,
                               std::vector<double> &ests) const {
    // eigen(Hermitian(H_c)): upper triangle, real diagonal
    const int c = completed;
    const bool full = completed == steps && completed < dim;
    std::vector<double> A((size_t)c * c), w, Vv;
    for (int i = 0; i < c; i++)
        for (int j = 0; j < c; j++) {
            int a = std::min(i, j), b = std::max(i, j);
            A[(size_t)i * c + j] = (i == j) ? h(i, i).real() : h(a, b).real();
        }
    sym_eig(c, A, w, Vv);
    coefs.assign((size_t)c * scales.size(), cd(0, 0));
    ests.assign(scales.size(), 0.0);
    std::vector<cd> g(c);
    for (size_t j = 0; j < scales.size(); j++) {
        if (scales[j] == cd(0, 0)) continue;
        for (int q = 0; q < c; q++) g[q] = std::exp(scales[j] * w[q]) * Vv[(size_t)0 * c + q];
        for (int i = 0; i < c; i++) {
            cd acc = 0;
            for (int q = 0; q < c; q++) acc += Vv[(size_t)i * c + q] * g[q];
            coefs[j * c + i] = beta * acc;
        }
        ests[j] = full ? std::abs(scales[j]) * terminal * std::abs(coefs[j * c + c - 1]) : 0.0;
    }
}

}  // namespace qs

namespace qs {

// lanczos_full / _lanczos_coefficients (tools.jl:2742-2850), step for step
// (queued with the scalars on the device; the stopping rule is applied
// afterwards in step order, and columns of Q past `used` are scratch):
// q = v0 / |v0|, residual = (A v0) / |v0|; for each index: store q, residual
// = A q - beta_prev q_prev (index > 1), alpha = Re <q, residual>, residual -=
// alpha q, optional sequential projection against every stored row, then
// beta = |residual| and q = residual / beta.  Q (Ns x m column-major, i.e.
// the reference's m x Ns Q_T row-major) is device memory of `dtype`.
void lanczos_device(OpHost &op, int dtype, const void *v0, int m, bool full_ortho, double eps, void *Q,
                    std::vector<double> &alphas, std::vector<double> &betas, int &used) {
    hipStream_t s = op_stream(op);
    const int64_t n = op.basis->Ns;
    const size_t esz = dtype_size(dtype);
    if (!(1 <= m && m < n)) fail(QS_EARG, "Lanczos dimension m must satisfy 1 <= m < length(v0)");
    VecWork wk(s);
    const double nrm = std::sqrt(vec_norm2(dtype, v0, n, wk, s));
    if (!(nrm > 0)) fail(QS_EARG, "initial Lanczos vector must be nonzero");
    const double tol = eps > 0 ? eps : ((dtype == QS_F32 || dtype == QS_C64) ? 1.1920928955078125e-07 : kEps);
    auto col = [&](int j) { return (void *)((char *)Q + (size_t)j * n * esz); };
    SBuf<char> res((size_t)n * esz, s);
    matvec(op, dtype, v0, res.get());
    vec_scale_copy(dtype, res.get(), res.get(), n, cd(1.0 / nrm, 0), s);
    vec_scale_copy(dtype, v0, col(0), n, cd(1.0 / nrm, 0), s);
    alphas.assign(m, 0.0);
    betas.assign(std::max(0, m - 1), 0.0);
    used = 0;
    // every step queued with its scalars in device memory: per index the
    // alpha dot, |res|^2 and (full_ortho) the m projection dots
    const size_t per = (size_t)2 * (m + 2);
    SBuf<double> d(per * m, s);
    vec_reserve(wk, n, 1);
    for (int index = 0; index < m; index++) {
        double *ha = d.get() + per * index, *n2 = ha + 2, *hp = ha + 4;
        void *q = col(index);
        if (index > 0) {
            matvec(op, dtype, q, res.get());
            vec_axpby_dev(dtype, d.get() + per * (index - 1) + 2, CO_NEG_SQRT, col(index - 1), cd(1, 0), res.get(),
                          n, s);
        }
        gram_dots_dev(dtype, q, n, 1, res.get(), wk, ha, s);
        vec_axpby_dev(dtype, ha, CO_NEG_RE, q, cd(1, 0), res.get(), n, s);
        if (full_ortho) {
            for (int p = 0; p <= index; p++) {
                gram_dots_dev(dtype, col(p), n, 1, res.get(), wk, hp + 2 * p, s);
                vec_axpby_dev(dtype, hp + 2 * p, CO_NEG, col(p), cd(1, 0), res.get(), n, s);
            }
        }
        if (index + 1 == m) break;
        vec_norm2_dev(dtype, res.get(), n, wk, n2, s);
        vec_axpby_dev(dtype, n2, CO_INV_SQRT, res.get(), cd(0, 0), col(index + 1), n, s);
    }
    std::vector<double> host(per * m);
    QS_HIP(hipMemcpyAsync(host.data(), d.get(), sizeof(double) * host.size(), hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    // the reference's stopping rule, in step order
    for (int index = 0; index < m; index++) {
        used = index + 1;
        alphas[index] = host[per * index];
        if (index + 1 == m) break;
        const double beta = std::sqrt(host[per * index + 2]);
        if (beta <= tol) break;
        betas[index] = beta;
    }
    alphas.resize(used);
    betas.resize(std::max(0, used - 1));
}

}  // namespace qs
