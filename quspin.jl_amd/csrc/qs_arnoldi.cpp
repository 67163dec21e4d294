// qs_arnoldi.cpp -- eigsh of a non-Hermitian operator on the device.
//
// The reference hands every QuantumLinearOperator to Arpack.eigs
// (operators.jl:3383-3390); for a non-Hermitian one that is ARPACK's complex
// implicitly restarted Arnoldi (znaupd) with which = SR / LR / SM / LM and its
// convergence test |r_i| <= tol * max(eps^(2/3), |theta_i|).  Here the same
// problem is solved by the Krylov-Schur restart (G. W. Stewart, SIAM J.
// Matrix Anal. Appl. 23 (2001) 601): an Arnoldi decomposition
// A V_m = V_m S + v_{m+1} b^H with the basis resident in HBM (two-pass
// classical Gram-Schmidt, FP64 accumulation, scalars kept on the device for a
// whole cycle), the small m x m matrix S brought to complex Schur form on the
// host, the wanted Ritz values moved to the top by Givens swaps, and the
// decomposition truncated to them.  Same answers as ARPACK's restart (both
// keep the wanted invariant subspace), fewer host round trips.
//
// Dense complex helpers (Householder Hessenberg reduction, single-shift QR,
// Schur reordering, triangular eigenvectors) follow the textbook algorithms
// (Golub & Van Loan, Matrix Computations, 4th ed., 7.4-7.6).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "qs_internal.h"

namespace qs {

namespace {

const double kEps = 2.220446049250313e-16;

struct Mat {   // row-major dense complex matrix
    int r = 0, c = 0;
    std::vector<cd> a;
    Mat() = default;
    Mat(int rows, int cols) : r(rows), c(cols), a((size_t)rows * cols, cd(0, 0)) {}
    cd &operator()(int i, int j) { return a[(size_t)i * c + j]; }
    cd operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

// unitary G = [[c, s], [-conj(s), c]], c real, with G [x; y] = [r; 0]
void givens(cd x, cd y, double &c, cd &s) {
    const double ax = std::abs(x), ay = std::abs(y);
    if (ay == 0) {
        c = 1;
        s = 0;
        return;
    }
    if (ax == 0) {
        c = 0;
        s = 1;
        return;
    }
    const double rho = std::hypot(ax, ay);
    c = ax / rho;
    s = (x / ax) * std::conj(y) / rho;
}
// rows i, j of M (columns c0 .. c1-1) <- G [row i; row j]
void rot_rows(Mat &M, int i, int j, double c, cd s, int c0, int c1) {
    for (int k = c0; k < c1; k++) {
        const cd a = M(i, k), b = M(j, k);
        M(i, k) = c * a + s * b;
        M(j, k) = -std::conj(s) * a + c * b;
    }
}
// columns i, j of M (rows r0 .. r1-1) <- [col i, col j] G^H
void rot_cols(Mat &M, int i, int j, double c, cd s, int r0, int r1) {
    for (int k = r0; k < r1; k++) {
        const cd a = M(k, i), b = M(k, j);
        M(k, i) = c * a + std::conj(s) * b;
        M(k, j) = -s * a + c * b;
    }
}

// A (n x n) -> upper Hessenberg H = Q^H A Q in place; Q accumulated
void hessenberg(Mat &A, Mat &Q) {
    const int n = A.r;
    Q = Mat(n, n);
    for (int i = 0; i < n; i++) Q(i, i) = 1;
    std::vector<cd> v(n);
    for (int k = 0; k + 2 < n; k++) {
        double xn = 0;
        for (int i = k + 1; i < n; i++) xn += std::norm(A(i, k));
        xn = std::sqrt(xn);
        if (xn == 0) continue;
        const cd x0 = A(k + 1, k);
        const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1, 0);
        const cd alpha = -ph * xn;
        for (int i = 0; i < n; i++) v[i] = 0;
        for (int i = k + 1; i < n; i++) v[i] = A(i, k);
        v[k + 1] -= alpha;
        double vn = 0;
        for (int i = k + 1; i < n; i++) vn += std::norm(v[i]);
        vn = std::sqrt(vn);
        if (vn == 0) continue;
        for (int i = k + 1; i < n; i++) v[i] /= vn;
        // A <- (I - 2 v v^H) A
        for (int j = 0; j < n; j++) {
            cd d = 0;
            for (int i = k + 1; i < n; i++) d += std::conj(v[i]) * A(i, j);
            for (int i = k + 1; i < n; i++) A(i, j) -= 2.0 * v[i] * d;
        }
        // A <- A (I - 2 v v^H), Q <- Q (I - 2 v v^H)
        for (Mat *M : {&A, &Q})
            for (int i = 0; i < n; i++) {
                cd d = 0;
                for (int j = k + 1; j < n; j++) d += (*M)(i, j) * v[j];
                for (int j = k + 1; j < n; j++) (*M)(i, j) -= 2.0 * d * std::conj(v[j]);
            }
        A(k + 1, k) = alpha;
        for (int i = k + 2; i < n; i++) A(i, k) = 0;
    }
}

// complex Schur form of a Hessenberg matrix: H <- T = Z^H H Z (upper
// triangular), Z <- Z * (accumulated rotations).  Wilkinson shifts with an
// exceptional shift every 10 stalled sweeps.
void schur_hessenberg(Mat &H, Mat &Z) {
    const int n = H.r;
    double hnorm = 0;
    for (cd z : H.a) hnorm = std::max(hnorm, std::abs(z));
    if (hnorm == 0) return;
    int hi = n - 1, stall = 0, total = 0;
    while (hi > 0) {
        int l = hi;
        for (; l > 0; l--) {
            double sc = std::abs(H(l - 1, l - 1)) + std::abs(H(l, l));
            if (sc == 0) sc = hnorm;
            if (std::abs(H(l, l - 1)) <= kEps * sc) {
                H(l, l - 1) = 0;
                break;
            }
        }
        if (l == hi) {
            hi--;
            stall = 0;
            continue;
        }
        if (++total > 60 * n) fail(QS_ENOCONV, "eigsh: the projected Schur iteration did not converge");
        cd mu;
        stall++;
        if (stall % 10 == 0) {
            mu = H(hi, hi) + std::abs(H(hi, hi - 1)) * 0.75;
        } else {
            const cd a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
            const cd half = 0.5 * (a - d);
            const cd disc = std::sqrt(half * half + b * c);
            const cd m1 = 0.5 * (a + d) + disc, m2 = 0.5 * (a + d) - disc;
            mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
        }
        // one implicit single-shift QR sweep over rows / columns l .. hi
        for (int k = l; k < hi; k++) {
            cd x, y;
            if (k == l) {
                x = H(l, l) - mu;
                y = H(l + 1, l);
            } else {
                x = H(k, k - 1);
                y = H(k + 1, k - 1);
            }
            double c;
            cd s;
            givens(x, y, c, s);
            rot_rows(H, k, k + 1, c, s, k == l ? l : k - 1, n);
            rot_cols(H, k, k + 1, c, s, 0, std::min(k + 3, hi + 1));
            rot_cols(Z, k, k + 1, c, s, 0, n);
            if (k > l) H(k + 1, k - 1) = 0;
        }
    }
    for (int i = 1; i < n; i++)
        for (int j = 0; j < i; j++) H(i, j) = 0;
}

// swap the diagonal entries i, i+1 of the upper triangular T (Z follows)
void schur_swap(Mat &T, Mat &Z, int i) {
    const int n = T.r;
    const cd t11 = T(i, i), t22 = T(i + 1, i + 1);
    double c;
    cd s;
    givens(T(i, i + 1), t22 - t11, c, s);
    rot_rows(T, i, i + 1, c, s, i + 2, n);
    rot_cols(T, i, i + 1, c, s, 0, i);
    rot_cols(Z, i, i + 1, c, s, 0, n);
    T(i, i) = t22;
    T(i + 1, i + 1) = t11;
}

// eigenvector of the upper triangular T for its diagonal entry i (unit norm)
std::vector<cd> tri_eigvec(const Mat &T, int i) {
    const int n = T.r;
    double tn = 0;
    for (cd z : T.a) tn = std::max(tn, std::abs(z));
    const double smin = std::max(kEps * tn, 1e-300);
    std::vector<cd> y(n, cd(0, 0));
    y[i] = 1;
    const cd lam = T(i, i);
    for (int j = i - 1; j >= 0; j--) {
        cd acc = 0;
        for (int l = j + 1; l <= i; l++) acc += T(j, l) * y[l];
        cd den = T(j, j) - lam;
        if (std::abs(den) < smin) den = smin;
        y[j] = -acc / den;
    }
    double nn = 0;
    for (cd z : y) nn += std::norm(z);
    nn = std::sqrt(nn);
    for (cd &z : y) z /= nn;
    return y;
}

// Schur form of a general n x n matrix: A = Z T Z^H
void schur_general(Mat &A, Mat &Z) {
    hessenberg(A, Z);
    schur_hessenberg(A, Z);
}

// operators.jl:3393-3401 orders (SR for SA, LR for LA, |.| for SM / LM)
bool more_wanted(int which, cd a, cd b) {
    switch (which) {
    case QS_WHICH_LA: return a.real() > b.real();
    case QS_WHICH_SM: return std::abs(a) < std::abs(b);
    case QS_WHICH_LM: return std::abs(a) > std::abs(b);
    default: return a.real() < b.real();
    }
}

std::vector<int> rank_wanted(const std::vector<cd> &theta, int which) {
    std::vector<int> idx(theta.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return more_wanted(which, theta[x], theta[y]); });
    return idx;
}

// move the diagonal entries listed in `order` (current positions) to the
// top of T, in that order
void schur_reorder(Mat &T, Mat &Z, const std::vector<int> &order) {
    const int n = T.r;
    std::vector<int> at(n);   // at[p] = original index of the entry now at p
    std::iota(at.begin(), at.end(), 0);
    for (size_t p = 0; p < order.size(); p++) {
        int cur = (int)(std::find(at.begin(), at.end(), order[p]) - at.begin());
        for (int i = cur - 1; i >= (int)p; i--) {
            schur_swap(T, Z, i);
            std::swap(at[i], at[i + 1]);
        }
    }
}

}  // namespace

void gen_eig(int n, const std::vector<cd> &A, std::vector<cd> &w, std::vector<cd> &V) {
    Mat M(n, n), Z;
    M.a = A;
    schur_general(M, Z);
    w.resize(n);
    V.assign((size_t)n * n, cd(0, 0));
    for (int i = 0; i < n; i++) {
        w[i] = M(i, i);
        std::vector<cd> y = tri_eigvec(M, i);
        for (int r = 0; r < n; r++) {
            cd acc = 0;
            for (int j = 0; j <= i; j++) acc += Z(r, j) * y[j];
            V[(size_t)r * n + i] = acc;
        }
    }
}

void eigs_device(OpHost &op, int dtype, int k, int which, double tol, int maxiter, int ncv, const void *v0_dev,
                 EigshResult &res, void *vecs_dev) {
    const int64_t n = op.basis->Ns;
    hipStream_t s = op_stream(op);
    if (!dtype_complex(dtype)) fail(QS_EARG, "a non-Hermitian operator needs a complex vector dtype");
    if (which == QS_WHICH_BE) fail(QS_EARG, "which=:BE only possible for real symmetric problem");
    if (tol <= 0) tol = kEps;
    if (maxiter <= 0) maxiter = 300;
    if (ncv <= 0) ncv = std::max(20, 2 * k + 1);
    ncv = (int)std::min<int64_t>(ncv, n);
    if (ncv <= k + 1) fail(QS_EARG, "ncv must exceed k + 1");
    const double eps23 = std::pow(kEps, 2.0 / 3.0);
    const size_t esz = dtype_size(dtype);
    const int m = ncv;
    SBuf<char> V((size_t)n * (m + 1) * esz, s);
    auto col = [&](int j) { return (void *)(V.get() + (size_t)j * n * esz); };
    VecWork wk(s);
    vec_reserve(wk, n, m + 1);
    if (v0_dev) QS_HIP(hipMemcpyAsync(col(0), v0_dev, (size_t)n * esz, hipMemcpyDeviceToDevice, s));
    else vec_random(dtype, col(0), n, 0x51a7ULL, s);
    const double nrm = std::sqrt(vec_norm2(dtype, col(0), n, wk, s));
    if (!(nrm > 0)) fail(QS_EARG, "initial vector must be nonzero");
    vec_scale_copy(dtype, col(0), col(0), n, cd(1.0 / nrm, 0), s);

    // B: rows 0..m-1 = S, row m = b^H (coefficients of v_{m+1})
    Mat B(m + 1, m);
    const size_t per = (size_t)4 * (m + 1) + 4;
    SBuf<double> d(per * m, s);
    std::vector<double> host(per * m);
    uint64_t reseed = 1;
    int kept = 0, nconv = 0;
    Mat T, Z;
    std::vector<cd> theta;
    std::vector<int> rank;
    auto orthonormal_random = [&](int j) {
        // column j + 1 <- random direction orthogonal to columns 0..j
        void *w = col(j + 1);
        vec_random(dtype, w, n, 0x9e37ULL + (reseed++), s);
        for (int pass = 0; pass < 2; pass++) {
            std::vector<cd> h;
            gram_dots(dtype, V.get(), n, j + 1, w, wk, s, h);
            gram_update(dtype, V.get(), n, j + 1, h, w, wk, s);
        }
        const double b2 = std::sqrt(vec_norm2(dtype, w, n, wk, s));
        vec_scale_copy(dtype, w, w, n, cd(1.0 / b2, 0), s);
    };
    for (int iter = 0; iter < maxiter; iter++) {
        int j0 = kept;
        while (j0 < m) {
            for (int j = j0; j < m; j++) {
                void *w = col(j + 1);
                double *h1 = d.get() + per * j, *h2 = h1 + 2 * (m + 1), *n2 = h2 + 2 * (m + 1), *n2a = n2 + 2;
                launch_matvec(op, dtype, col(j), w, 1.0, 0.0, 0.0, 0.0, false, false, false, s);
                // znaitr's DGKS step: a second pass only if |w'| <= 0.717 |w|
                gram_dgks_dev(dtype, V.get(), n, j + 1, w, wk, h1, h2, n2a, n2, s);
                vec_axpby_dev(dtype, n2, CO_INV_SQRT, w, cd(0, 0), w, n, s);
            }
            QS_HIP(hipMemcpyAsync(host.data() + per * j0, d.get() + per * j0, sizeof(double) * per * (m - j0),
                                  hipMemcpyDeviceToHost, s));
            QS_HIP(hipStreamSynchronize(s));
            int next = m;
            for (int j = j0; j < m; j++) {
                res.nmatvec++;
                const double *h1 = host.data() + per * j, *h2 = h1 + 2 * (m + 1);
                for (int i = 0; i <= j; i++) B(i, j) = cd(h1[2 * i], h1[2 * i + 1]) + cd(h2[2 * i], h2[2 * i + 1]);
                double snorm = 0;
                for (cd z : B.a) snorm = std::max(snorm, std::abs(z));
                double beta = std::sqrt(h2[2 * (m + 1)]);
                const bool brk = beta <= 1e3 * kEps * std::max(1.0, snorm);
                if (brk) {
                    // invariant subspace: continue from a fresh direction
                    beta = 0;
                    orthonormal_random(j);
                }
                B(j + 1, j) = beta;
                if (brk) {
                    next = j + 1;
                    break;
                }
            }
            j0 = next;
        }
        // Schur form of S and the Ritz pairs
        T = Mat(m, m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) T(i, j) = B(i, j);
        schur_general(T, Z);
        theta.resize(m);
        for (int i = 0; i < m; i++) theta[i] = T(i, i);
        rank = rank_wanted(theta, which);
        // the wanted k first, so their eigenvectors of T are local
        schur_reorder(T, Z, std::vector<int>(rank.begin(), rank.begin() + k));
        for (int i = 0; i < m; i++) theta[i] = T(i, i);
        // residuals |b^H Z y_i| of the k wanted (now at the top)
        std::vector<cd> bz(m, cd(0, 0));
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) bz[j] += B(m, i) * Z(i, j);
        nconv = 0;
        for (int i = 0; i < k; i++) {
            std::vector<cd> y = tri_eigvec(T, i);
            cd r = 0;
            for (int j = 0; j <= i; j++) r += bz[j] * y[j];
            if (std::abs(r) <= tol * std::max(eps23, std::abs(theta[i]))) nconv++;
        }
        if (nconv >= k || iter == maxiter - 1) break;
        // Krylov-Schur restart: keep the `keep` most wanted Schur vectors
        const int keep = std::min(m - 1, k + (m - k) / 2);
        rank = rank_wanted(theta, which);
        schur_reorder(T, Z, std::vector<int>(rank.begin(), rank.begin() + keep));
        std::vector<cd> C((size_t)m * keep);
        for (int c = 0; c < keep; c++)
            for (int j = 0; j < m; j++) C[(size_t)c * m + j] = Z(j, c);
        {
            SBuf<char> tmp((size_t)n * keep * esz, s);
            basis_combine(dtype, V.get(), n, m, C, keep, tmp.get(), s);
            QS_HIP(hipMemcpyAsync(col(0), tmp.get(), (size_t)n * keep * esz, hipMemcpyDeviceToDevice, s));
        }
        QS_HIP(hipMemcpyAsync(col(keep), col(m), (size_t)n * esz, hipMemcpyDeviceToDevice, s));
        std::vector<cd> bnew(keep, cd(0, 0));
        for (int c = 0; c < keep; c++)
            for (int i = 0; i < m; i++) bnew[c] += B(m, i) * Z(i, c);
        B = Mat(m + 1, m);
        for (int i = 0; i < keep; i++)
            for (int j = i; j < keep; j++) B(i, j) = T(i, j);
        for (int c = 0; c < keep; c++) B(keep, c) = bnew[c];
        kept = keep;
    }
    // T, Z hold the wanted k at the top; order them as operators.jl:3393-3401
    std::vector<int> out(k);
    std::iota(out.begin(), out.end(), 0);
    std::stable_sort(out.begin(), out.end(), [&](int a, int b) { return more_wanted(which, theta[a], theta[b]); });
    res.vals.resize(2 * (size_t)k);
    for (int i = 0; i < k; i++) {
        res.vals[2 * i] = theta[out[i]].real();
        res.vals[2 * i + 1] = theta[out[i]].imag();
    }
    res.nconv = nconv;
    if (vecs_dev) {
        std::vector<cd> C((size_t)m * k);
        for (int c = 0; c < k; c++) {
            std::vector<cd> y = tri_eigvec(T, out[c]);
            for (int j = 0; j < m; j++) {
                cd acc = 0;
                for (int l = 0; l <= out[c]; l++) acc += Z(j, l) * y[l];
                C[(size_t)c * m + j] = acc;
            }
        }
        basis_combine(dtype, V.get(), n, m, C, k, vecs_dev, s);
    }
    QS_HIP(hipStreamSynchronize(s));
}

}  // namespace qs
