// qs_rk45.hip -- time evolution under a time-dependent Hamiltonian
// H(t) = H_0 + sum_j f_j(t) H_j on the device.
//
// Reference: evolve(::Hamiltonian, v0, t0, times; imag_time, max_step, rtol,
// atol) with dynamic terms (operators.jl:2092-2128) and its adaptive
// Dormand-Prince step _rk45_interval (operators.jl:1964-2071).  The step
// control, the accept / reject rule and the error norm are the reference's;
// every product H(t) v runs on the device (the matrix-free kernels of each
// operator, one launch per term), the stage combinations are one fused
// launch each, and one scalar -- the step's error norm -- comes back to the
// host per step.  The drive values f_j(t) come from a host callback, once per
// stage time, so they may be any function of t.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "qs_device.h"

namespace qs {

namespace {

constexpr int kRkBlock = 256;
constexpr int kRkMaxVec = 8;

// out = base + step * (c_0 v_0 + c_1 v_1 + ...), the sum taken left to right
struct RkComb {
    const void *v[kRkMaxVec];
    double c[kRkMaxVec];
    int nv;
};

template <typename V>
__global__ __launch_bounds__(kRkBlock) void k_rk_comb(const V *__restrict__ base, RkComb C, double step,
                                                      V *__restrict__ out, int64_t n) {
    using T = VTraits<V>;
    for (int64_t i = (int64_t)blockIdx.x * kRkBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRkBlock) {
        double sr = 0, si = 0;
        for (int j = 0; j < C.nv; j++) {
            double vr, vi;
            T::load((const V *)C.v[j], i, vr, vi);
            sr += C.c[j] * vr;
            si += C.c[j] * vi;
        }
        double br, bi;
        T::load(base, i, br, bi);
        T::store(out, i, br + step * sr, bi + step * si);
    }
}

// partial[b] = max over the block's rows of |fifth - fourth| / (atol + rtol max(|state|, |fifth|))
template <typename V>
__global__ __launch_bounds__(kRkBlock) void k_rk_err(const V *__restrict__ state, const V *__restrict__ fifth,
                                                     const V *__restrict__ fourth, double atol, double rtol, int64_t n,
                                                     double *__restrict__ partial) {
    using T = VTraits<V>;
    __shared__ double sh[kRkBlock];
    double e = 0;
    for (int64_t i = (int64_t)blockIdx.x * kRkBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRkBlock) {
        double sr, si, fr, fi, qr, qi;
        T::load(state, i, sr, si);
        T::load(fifth, i, fr, fi);
        T::load(fourth, i, qr, qi);
        const double scale = atol + rtol * fmax(hypot(sr, si), hypot(fr, fi));
        e = fmax(e, hypot(fr - qr, fi - qi) / scale);
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = kRkBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kRkBlock) void k_rk_max(const double *__restrict__ partial, int nb, double *out) {
    __shared__ double sh[kRkBlock];
    double e = 0;
    for (int i = threadIdx.x; i < nb; i += kRkBlock) e = fmax(e, partial[i]);
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = kRkBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

int rk_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kRkBlock - 1) / kRkBlock, 4096)); }

// spacing of x, Julia's eps(x) for x >= 0
double jeps(double x) {
    x = std::fabs(x);
    return std::nextafter(x, std::numeric_limits<double>::infinity()) - x;
}

struct RkState {
    const OpHost *h0;
    std::vector<const OpHost *> dyn;
    qs_drive_fn drive;
    void *user;
    int dtype;
    int64_t n;       // rows of one column
    int64_t ncols;
    bool imag;
    hipStream_t s;
    std::vector<double> coef;   // 2 per dynamic term
    int64_t nmatvec = 0;
};

// out = (-i or -1) (H_0 + sum_j f_j(t) H_j) x, column by column
void derivative(RkState &R, const void *x, void *out, double t) {
    const size_t esz = dtype_size(R.dtype);
    if (!R.dyn.empty()) {
        R.coef.assign(2 * R.dyn.size(), 0.0);
        R.drive(t, (int)R.dyn.size(), R.coef.data(), R.user);
    }
    const cd s = R.imag ? cd(-1, 0) : cd(0, -1);
    for (int64_t c = 0; c < R.ncols; c++) {
        const char *xc = (const char *)x + (size_t)c * R.n * esz;
        char *yc = (char *)out + (size_t)c * R.n * esz;
        launch_matvec(*R.h0, R.dtype, xc, yc, s.real(), s.imag(), 0.0, 0.0, false, false, false, R.s);
        R.nmatvec++;
        for (size_t j = 0; j < R.dyn.size(); j++) {
            const cd f(R.coef[2 * j], R.coef[2 * j + 1]);
            if (f == cd(0, 0)) continue;   // iszero(coefficient) && continue
            const cd a = s * f;
            launch_matvec(*R.dyn[j], R.dtype, xc, yc, a.real(), a.imag(), 1.0, 0.0, false, false, false, R.s);
            R.nmatvec++;
        }
    }
}

template <typename V>
void comb(RkState &R, const void *base, std::initializer_list<std::pair<const void *, double>> terms, double step,
          void *out) {
    RkComb C;
    memset(&C, 0, sizeof C);
    for (auto &p : terms) {
        C.v[C.nv] = p.first;
        C.c[C.nv] = p.second;
        C.nv++;
    }
    const int64_t n = R.n * R.ncols;
    hipLaunchKernelGGL((k_rk_comb<V>), dim3(rk_grid(n)), dim3(kRkBlock), 0, R.s, (const V *)base, C, step, (V *)out, n);
    QS_HIP(hipGetLastError());
}

// _rk45_interval (operators.jl:1964-2071) from start to target; state is
// updated in place, the returned value is the next step size.
template <typename V>
double rk45_interval(RkState &R, void *state, double start, double target, double initial_step, double max_step,
                     double rtol, double atol, SBuf<char> &work, SBuf<double> &red, int64_t &steps) {
    const size_t bytes = (size_t)R.n * R.ncols * sizeof(V);
    auto vec = [&](int i) { return (void *)(work.get() + (size_t)i * bytes); };
    void *k[8] = {nullptr, vec(0), vec(1), vec(2), vec(3), vec(4), vec(5), vec(6)};
    void *tmp = vec(7), *fifth = vec(8), *fourth = vec(9);
    const int64_t n = R.n * R.ncols;
    const int nb = rk_grid(n);
    double time = start;
    double step = std::min(max_step, std::max(initial_step, jeps(target + 1.0)));
    while (time < target) {
        const double remaining = target - time;
        if (remaining <= 8 * jeps(std::max(std::max(std::fabs(time), std::fabs(target)), 1.0))) {
            time = target;
            break;
        }
        step = std::min(step, remaining);
        if (!(step > jeps(std::max(std::fabs(time), 1.0)))) fail(QS_ENOCONV, "adaptive evolution step underflow");
        const double h = step;
        derivative(R, state, k[1], time);
        comb<V>(R, state, {{k[1], 1.0 / 5}}, h, tmp);
        derivative(R, tmp, k[2], time + h * (1.0 / 5));
        comb<V>(R, state, {{k[1], 3.0 / 40}, {k[2], 9.0 / 40}}, h, tmp);
        derivative(R, tmp, k[3], time + h * (3.0 / 10));
        comb<V>(R, state, {{k[1], 44.0 / 45}, {k[2], -56.0 / 15}, {k[3], 32.0 / 9}}, h, tmp);
        derivative(R, tmp, k[4], time + h * (4.0 / 5));
        comb<V>(R, state,
                {{k[1], 19372.0 / 6561}, {k[2], -25360.0 / 2187}, {k[3], 64448.0 / 6561}, {k[4], -212.0 / 729}}, h,
                tmp);
        derivative(R, tmp, k[5], time + h * (8.0 / 9));
        comb<V>(R, state,
                {{k[1], 9017.0 / 3168}, {k[2], -355.0 / 33}, {k[3], 46732.0 / 5247}, {k[4], 49.0 / 176},
                 {k[5], -5103.0 / 18656}},
                h, tmp);
        derivative(R, tmp, k[6], time + h);
        comb<V>(R, state,
                {{k[1], 35.0 / 384}, {k[3], 500.0 / 1113}, {k[4], 125.0 / 192}, {k[5], -2187.0 / 6784},
                 {k[6], 11.0 / 84}},
                h, fifth);
        derivative(R, fifth, k[7], time + h);
        comb<V>(R, state,
                {{k[1], 5179.0 / 57600}, {k[3], 7571.0 / 16695}, {k[4], 393.0 / 640}, {k[5], -92097.0 / 339200},
                 {k[6], 187.0 / 2100}, {k[7], 1.0 / 40}},
                h, fourth);
        hipLaunchKernelGGL((k_rk_err<V>), dim3(nb), dim3(kRkBlock), 0, R.s, (const V *)state, (const V *)fifth,
                           (const V *)fourth, atol, rtol, n, red.get());
        hipLaunchKernelGGL(k_rk_max, dim3(1), dim3(kRkBlock), 0, R.s, red.get(), nb, red.get() + nb);
        QS_HIP(hipGetLastError());
        double error = 0;
        QS_HIP(hipMemcpyAsync(&error, red.get() + nb, sizeof(double), hipMemcpyDeviceToHost, R.s));
        QS_HIP(hipStreamSynchronize(R.s));
        steps++;
        if (error <= 1) {
            QS_HIP(hipMemcpyAsync(state, fifth, bytes, hipMemcpyDeviceToDevice, R.s));
            time += h;
        }
        const double factor = error == 0 ? 5.0 : std::min(5.0, std::max(0.2, 0.9 * std::pow(error, -1.0 / 5)));
        step = std::min(max_step, h * factor);
    }
    return step;
}

template <typename V>
void evolve_dynamic_t(RkState &R, const void *v0, double t0, const std::vector<double> &times, double rtol,
                      double atol, double max_step, void *out, int64_t &steps) {
    const size_t bytes = (size_t)R.n * R.ncols * sizeof(V);
    SBuf<char> state(bytes, R.s), work(bytes * 10, R.s);
    const int nb = rk_grid(R.n * R.ncols);
    SBuf<double> red((size_t)nb + 1, R.s);
    QS_HIP(hipMemcpyAsync(state.get(), v0, bytes, hipMemcpyHostToDevice, R.s));
    double current = t0, next_step = max_step;
    VecWork wk(R.s);
    for (size_t it = 0; it < times.size(); it++) {
        const double target = times[it];
        if (!(target >= current)) fail(QS_EARG, "times must be sorted and not precede t0");
        next_step = rk45_interval<V>(R, state.get(), current, target, next_step, max_step, rtol, atol, work, red,
                                     steps);
        current = target;
        if (R.imag) {
            // state ./ sqrt.(sum(abs2, state; dims=1))
            for (int64_t c = 0; c < R.ncols; c++) {
                void *col = state.get() + (size_t)c * R.n * sizeof(V);
                const double nrm2 = vec_norm2(R.dtype, col, R.n, wk, R.s);
                vec_scale_copy(R.dtype, col, col, R.n, cd(1.0 / std::sqrt(nrm2), 0), R.s);
            }
        }
        QS_HIP(hipMemcpyAsync((char *)out + it * bytes, state.get(), bytes, hipMemcpyDeviceToHost, R.s));
    }
    QS_HIP(hipStreamSynchronize(R.s));
}

}  // namespace

void evolve_dynamic_host(const OpHost &h0, const std::vector<const OpHost *> &dyn, qs_drive_fn drive, void *user,
                         int dtype, const void *v0, int64_t ncols, double t0, const std::vector<double> &times,
                         int imag_time, double rtol, double atol, double max_step, void *out, int64_t *n_steps,
                         int64_t *n_matvec) {
    if (dtype != QS_C128 && dtype != QS_C64) fail(QS_EARG, "evolve needs a complex vector dtype");
    if (!(max_step > 0)) fail(QS_EARG, "max_step must be positive");
    if (ncols < 1) fail(QS_EDIM, "ncols must be positive");
    if (!dyn.empty() && !drive) fail(QS_EARG, "dynamic terms need a drive function");
    for (const OpHost *d : dyn)
        if (d->basis != h0.basis) fail(QS_EDIM, "dynamic terms must act on the static term's basis");
    RkState R{&h0, dyn, drive, user, dtype, h0.basis->Ns, ncols, imag_time != 0, op_stream(h0), {}, 0};
    int64_t steps = 0;
    if (dtype == QS_C128) evolve_dynamic_t<cplx_d>(R, v0, t0, times, rtol, atol, max_step, out, steps);
    else evolve_dynamic_t<cplx_f>(R, v0, t0, times, rtol, atol, max_step, out, steps);
    if (n_steps) *n_steps = steps;
    if (n_matvec) *n_matvec = R.nmatvec;
}

}  // namespace qs
