// qs_vec.hip -- Krylov vector kernels with FP64 accumulation.
//
// Every reduction is deterministic: a fixed number of blocks, each owning a
// fixed contiguous row range, a fixed-shape tree inside the block and a
// second fixed-shape tree over the block partials.  Results therefore do not
// depend on what else shares the device.  Vectors stored in FP32 (the
// FP32-stream form) are widened to FP64 before every product.
#include <type_traits>
#include <algorithm>

#include "qs_device.h"

namespace qs {

size_t dtype_size(int dtype) {
    switch (dtype) {
    case QS_F64: return 8;
    case QS_C128: return 16;
    case QS_F32: return 4;
    case QS_C64: return 8;
    }
    fail(QS_EARG, "unknown dtype");
}
bool dtype_complex(int dtype) { return dtype == QS_C128 || dtype == QS_C64; }

namespace {

constexpr int kVB = 256;
constexpr int kMaxBlocks = 2048;
constexpr int kCols = 8;   // Gram columns per pass
constexpr int kDgksFuse = 24;   // basis columns the fused update + Gram holds in registers

// butterfly over the 64 lanes of a wavefront: every lane ends with the same
// sum (each step adds the same two operands on both partners)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void block_reduce2(double &re, double &im, double *sh) {
    // fixed shape: wavefront butterflies, then the four wavefront sums in a
    // fixed order; every thread returns the block sum
    static_assert(kVB == 256, "four wavefronts per block");
    re = wave_sum(re);
    im = wave_sum(im);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[w] = re;
        sh[4 + w] = im;
    }
    __syncthreads();
    re = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    im = (sh[4] + sh[5]) + (sh[6] + sh[7]);
    __syncthreads();
}

// DGKS test of ARPACK's Arnoldi step (dsaitr / znaitr, eta = 0.717): after
// the first Gram-Schmidt pass w' = w - V h1 the second pass is needed only if
// |w'| <= eta |w|, with |w|^2 = |w'|^2 + |h1|^2.  Read from device memory by
// every block of the second-pass launches, so the decision costs no host
// round trip; every block computes it in the same order.
__device__ __forceinline__ bool dgks_skip(const double *h1, int m, const double *n2w) {
    double hh = 0;
    for (int j = 0; j < m; j++) hh += h1[2 * j] * h1[2 * j] + h1[2 * j + 1] * h1[2 * j + 1];
    const double e2 = 0.717 * 0.717;
    return n2w[0] > e2 * (n2w[0] + hh);
}

// Block b of nb owns one contiguous row range.  Kernel bodies take (b, nb)
// explicitly so a batched launch (many jobs, one grid) gives every job the
// partition -- and therefore the sums -- of its own single launch.
__device__ __forceinline__ void row_range(int64_t n, int b, int nb, int64_t &r0, int64_t &r1) {
    int64_t chunk = (n + nb - 1) / nb;
    r0 = (int64_t)b * chunk;
    r1 = r0 + chunk < n ? r0 + chunk : n;
}

// Column groups of NC (<= 8) columns: every load of a row pair is issued
// before any is used (the second row of the pair clamped to the range, its
// terms dropped), so each lane keeps 2 (NC + 1) loads in flight.
template <typename V, int NC>
__device__ __forceinline__ void gram_cols(const V *__restrict__ Vb, int64_t n, int j0, const V *__restrict__ w,
                                          int64_t r0, int64_t r1, double (&acc_r)[8], double (&acc_i)[8]) {
    using T = VTraits<V>;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 2 * kVB) {
        const int64_t rb = r + kVB < r1 ? r + kVB : r1 - 1;
        const bool hb = r + kVB < r1;
        V wa = w[r], wb = w[rb];
        V va[NC], vb[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            va[c] = Vb[(int64_t)(j0 + c) * n + r];
            vb[c] = Vb[(int64_t)(j0 + c) * n + rb];
        }
        double war, wai, wbr, wbi;
        T::load(&wa, 0, war, wai);
        T::load(&wb, 0, wbr, wbi);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double vr, vi;
            T::load(&va[c], 0, vr, vi);
            acc_r[c] += vr * war + vi * wai;
            acc_i[c] += vr * wai - vi * war;
            if (hb) {
                T::load(&vb[c], 0, vr, vi);
                acc_r[c] += vr * wbr + vi * wbi;
                acc_i[c] += vr * wbi - vi * wbr;
            }
        }
    }
}

// partial[(j * nb + b) * 2 + {0,1}] = sum over block b rows of conj(V[r, j]) * w[r]
template <typename V>
__device__ __forceinline__ void gram_body(const V *__restrict__ Vb, int64_t n, int m, const V *__restrict__ w,
                                          double *partial, int b, int nb, double *sh) {
    int64_t r0, r1;
    row_range(n, b, nb, r0, r1);
    for (int j0 = 0; j0 < m; j0 += kCols) {
        double acc_r[kCols], acc_i[kCols];
#pragma unroll
        for (int c = 0; c < kCols; c++) { acc_r[c] = 0; acc_i[c] = 0; }
        const int nc = m - j0 < kCols ? m - j0 : kCols;
        switch (nc) {
        case 1: gram_cols<V, 1>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        case 2: gram_cols<V, 2>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        case 3: gram_cols<V, 3>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        case 4: gram_cols<V, 4>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        case 5: gram_cols<V, 5>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        case 6: gram_cols<V, 6>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        case 7: gram_cols<V, 7>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        default: gram_cols<V, 8>(Vb, n, j0, w, r0, r1, acc_r, acc_i); break;
        }
        for (int c = 0; c < nc; c++) {
            double re = acc_r[c], im = acc_i[c];
            block_reduce2(re, im, sh);
            if (threadIdx.x == 0) {
                partial[((int64_t)(j0 + c) * nb + b) * 2] = re;
                partial[((int64_t)(j0 + c) * nb + b) * 2 + 1] = im;
            }
        }
    }
}

// out[j] = sum_b partial[j][b] in a fixed tree (one block per j)
__device__ __forceinline__ void reduce_body(const double *partial, int nb, double *out, int j, double *sh) {
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nb; b += kVB) {
        re += partial[((int64_t)j * nb + b) * 2];
        im += partial[((int64_t)j * nb + b) * 2 + 1];
    }
    block_reduce2(re, im, sh);
    if (threadIdx.x == 0) {
        out[2 * j] = re;
        out[2 * j + 1] = im;
    }
}

// w[ra], w[rb] -= sum_j V[., j] h[j] for the rows' whole column list, in
// column order; the columns' loads for both rows go out eight at a time
// (clamped to column m - 1, their terms dropped) before any is used.
template <typename V>
__device__ __forceinline__ void update_rows(const V *__restrict__ Vb, int64_t n, int m, const double *h, int64_t ra,
                                            int64_t rb, double &war, double &wai, double &wbr, double &wbi) {
    using T = VTraits<V>;
    for (int j0 = 0; j0 < m; j0 += 8) {
        V va[8], vb[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int64_t col = (int64_t)(j0 + c < m ? j0 + c : m - 1) * n;
            va[c] = Vb[col + ra];
            vb[c] = Vb[col + rb];
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
            if (j0 + c < m) {
                const double hr = h[2 * (j0 + c)], hi = h[2 * (j0 + c) + 1];
                double vr, vi;
                T::load(&va[c], 0, vr, vi);
                war -= vr * hr - vi * hi;
                wai -= vr * hi + vi * hr;
                T::load(&vb[c], 0, vr, vi);
                wbr -= vr * hr - vi * hi;
                wbi -= vr * hi + vi * hr;
            }
        }
    }
}

// w[r] -= sum_j V[r, j] h[j]; partial[b] = sum |w[r]|^2 (of the stored values).
// Rows go in pairs (r, r + kVB), the second dropped past the range end.
template <typename V>
__device__ __forceinline__ void update_body(const V *__restrict__ Vb, int64_t n, int m, const double *h,
                                            V *__restrict__ w, double *partial, int b, int nb, double *sh) {
    using T = VTraits<V>;
    int64_t r0, r1;
    row_range(n, b, nb, r0, r1);
    double nrm = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 2 * kVB) {
        const bool hb = r + kVB < r1;
        const int64_t rb = hb ? r + kVB : r;
        double war, wai, wbr, wbi;
        T::load(w, r, war, wai);
        T::load(w, rb, wbr, wbi);
        if (m > 0) update_rows<V>(Vb, n, m, h, r, rb, war, wai, wbr, wbi);
        T::store(w, r, war, wai);
        T::load(w, r, war, wai);
        nrm += war * war + wai * wai;
        if (hb) {
            T::store(w, rb, wbr, wbi);
            T::load(w, rb, wbr, wbi);
            nrm += wbr * wbr + wbi * wbi;
        }
    }
    double dummy = 0;
    block_reduce2(nrm, dummy, sh);
    if (threadIdx.x == 0) {
        partial[2 * b] = nrm;
        partial[2 * b + 1] = 0;
    }
}

// The first pass's update fused with the second pass's Gram product (m <=
// MX columns, held in registers between the two): w' = w - V h1 (the same
// per-row arithmetic as update_body), partial[j] = conj(V_j) . w' for j < m
// and partial[m] = |w'|^2.  One read of V instead of two.
template <typename V, int MX>
__device__ __forceinline__ void update_gram_body(const V *__restrict__ Vb, int64_t n, int m, const double *h,
                                                 V *__restrict__ w, double *partial, int b, int nb, double *sh) {
    using T = VTraits<V>;
    int64_t r0, r1;
    row_range(n, b, nb, r0, r1);
    double gr[MX], gi[MX];
#pragma unroll
    for (int c = 0; c < MX; c++) { gr[c] = 0; gi[c] = 0; }
    double nrm = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kVB) {
        V vv[MX];
#pragma unroll
        for (int c = 0; c < MX; c++) vv[c] = Vb[(int64_t)(c < m ? c : m - 1) * n + r];
        double wr, wi;
        T::load(w, r, wr, wi);
#pragma unroll
        for (int c = 0; c < MX; c++) {
            if (c < m) {
                double vr, vi;
                T::load(&vv[c], 0, vr, vi);
                const double hr = h[2 * c], hi = h[2 * c + 1];
                wr -= vr * hr - vi * hi;
                wi -= vr * hi + vi * hr;
            }
        }
        T::store(w, r, wr, wi);
        T::load(w, r, wr, wi);
        nrm += wr * wr + wi * wi;
#pragma unroll
        for (int c = 0; c < MX; c++) {
            if (c < m) {
                double vr, vi;
                T::load(&vv[c], 0, vr, vi);
                gr[c] += vr * wr + vi * wi;
                gi[c] += vr * wi - vi * wr;
            }
        }
    }
    // unrolled with the column test inside, so gr / gi stay in registers
#pragma unroll
    for (int c = 0; c < MX; c++) {
        if (c < m) {   // uniform over the block
            double re = gr[c], im = gi[c];
            block_reduce2(re, im, sh);
            if (threadIdx.x == 0) {
                partial[((int64_t)c * nb + b) * 2] = re;
                partial[((int64_t)c * nb + b) * 2 + 1] = im;
            }
        }
    }
    double dummy = 0;
    block_reduce2(nrm, dummy, sh);
    if (threadIdx.x == 0) {
        partial[((int64_t)m * nb + b) * 2] = nrm;
        partial[((int64_t)m * nb + b) * 2 + 1] = 0;
    }
}

template <typename V>
__device__ __forceinline__ void norm2_body(const V *__restrict__ x, int64_t n, double *partial, int b, int nb,
                                           double *sh) {
    using T = VTraits<V>;
    int64_t r0, r1;
    row_range(n, b, nb, r0, r1);
    double nrm = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kVB) {
        double a, c;
        T::load(x, r, a, c);
        nrm += a * a + c * c;
    }
    double dummy = 0;
    block_reduce2(nrm, dummy, sh);
    if (threadIdx.x == 0) {
        partial[2 * b] = nrm;
        partial[2 * b + 1] = 0;
    }
}


template <typename V>
__device__ __forceinline__ void axpby_body(double ar, double ai, const V *__restrict__ x, double br, double bi,
                                           V *__restrict__ y, int64_t n, int b_zero, int b, int nb,
                                           const double *acoef = nullptr, int cmode = CO_HOST) {
    using T = VTraits<V>;
    if (acoef) dev_coef(acoef, cmode, ar, ai);
    for (int64_t r = (int64_t)b * blockDim.x + threadIdx.x; r < n; r += (int64_t)nb * blockDim.x) {
        double xr, xi, yr = 0, yi = 0;
        T::load(x, r, xr, xi);
        if (!b_zero) T::load(y, r, yr, yi);
        double orr = ar * xr - ai * xi + (b_zero ? 0.0 : br * yr - bi * yi);
        double oi = ar * xi + ai * xr + (b_zero ? 0.0 : br * yi + bi * yr);
        T::store(y, r, orr, oi);
    }
}

// Y[:, i] = sum_j V[:, j] C[j + i*m]   (C interleaved complex, column-major m x k)
// Outputs in groups of eight per pass over the row; each group reads the
// row's basis values eight columns at a time (clamped loads, all issued
// before use) and adds them in column order, so each output's sum has the
// order of a plain loop over j.
template <typename V>
__device__ __forceinline__ void combine_body(const V *__restrict__ Vb, int64_t n, int m, const double *C, int k,
                                             V *__restrict__ Y, int b, int nb) {
    using T = VTraits<V>;
    for (int64_t r = (int64_t)b * blockDim.x + threadIdx.x; r < n; r += (int64_t)nb * blockDim.x) {
        for (int i0 = 0; i0 < k; i0 += 8) {
            double ar[8], ai[8];
#pragma unroll
            for (int q = 0; q < 8; q++) { ar[q] = 0; ai[q] = 0; }
            for (int j0 = 0; j0 < m; j0 += 8) {
                V vv[8];
#pragma unroll
                for (int c = 0; c < 8; c++) vv[c] = Vb[(int64_t)(j0 + c < m ? j0 + c : m - 1) * n + r];
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    if (j0 + c >= m) break;
                    double vr, vi;
                    T::load(&vv[c], 0, vr, vi);
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        if (i0 + q < k) {
                            const int64_t e = (int64_t)(i0 + q) * m + j0 + c;
                            const double cr = C[2 * e], ci = C[2 * e + 1];
                            ar[q] += vr * cr - vi * ci;
                            ai[q] += vr * ci + vi * cr;
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (i0 + q < k) T::store(Y, (int64_t)(i0 + q) * n + r, ar[q], ai[q]);
        }
    }
}

// Y[:, :k] = V[:, :m] C on the matrix cores (v_mfma_f64_16x16x4f64): one
// wavefront computes a 16-row x 64-column tile of Y as four 16 x 16 MFMA
// accumulators, stepping through m four basis columns at a time.  Operand
// layout of the instruction (checked on gfx950): lane l holds A[l % 16][l / 16]
// of the 16 x 4 block of V and B[l / 16][l % 16] of the 4 x 16 block of C;
// accumulator register i of lane l is D[l / 16 + 4 i][l % 16].  Complex values run as
// four real products (re.re - im.im, re.im + im.re).  FP64 accumulation.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kMfmaCols = 64;

template <bool CPLX>
__global__ __launch_bounds__(256) void k_combine_mfma(const double *__restrict__ Vb, int64_t n, int m,
                                                      const double *__restrict__ C, int k, double *__restrict__ Y) {
    const int lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int col0 = blockIdx.y * kMfmaCols;
    if (row0 >= n) return;   // whole wavefront: no MFMA issued
    const int ar = lane & 15, ak = lane >> 4;            // A operand: row ar, basis column ak
    const int bk = lane >> 4, bc = lane & 15;            // B operand: basis column bk, output column bc
    const int64_t arow = row0 + ar;
    v4d accr[4], acci[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        accr[t] = v4d{0, 0, 0, 0};
        acci[t] = v4d{0, 0, 0, 0};
    }
    for (int j0 = 0; j0 < m; j0 += 4) {
        const int j = j0 + ak;
        double vr = 0, vi = 0;
        if (j < m && arow < n) {
            if (CPLX) {
                vr = Vb[2 * ((int64_t)j * n + arow)];
                vi = Vb[2 * ((int64_t)j * n + arow) + 1];
            } else {
                vr = Vb[(int64_t)j * n + arow];
            }
        }
        const int jb = j0 + bk;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int col = col0 + 16 * t + bc;
            double cr = 0, ci = 0;
            if (jb < m && col < k) {
                cr = C[2 * ((int64_t)col * m + jb)];
                ci = C[2 * ((int64_t)col * m + jb) + 1];
            }
            accr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(vr, cr, accr[t], 0, 0, 0);
            if (CPLX) {
                accr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(vi, -ci, accr[t], 0, 0, 0);
                acci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(vr, ci, acci[t], 0, 0, 0);
                acci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(vi, cr, acci[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = col0 + 16 * t + (lane & 15);
        if (col >= k) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t r = row0 + (lane >> 4) + 4 * i;
            if (r >= n) continue;
            if (CPLX) {
                Y[2 * ((int64_t)col * n + r)] = accr[t][i];
                Y[2 * ((int64_t)col * n + r) + 1] = acci[t][i];
            } else {
                Y[(int64_t)col * n + r] = accr[t][i];
            }
        }
    }
}

// The second-pass launches take (dh, dn, dm): when non-null and the DGKS
// test says no second pass is needed, the Gram and update kernels return at
// once, the Gram reduction writes zeros and the norm reduction copies |w'|^2.
template <typename V>
__global__ __launch_bounds__(kVB) void k_gram(const V *__restrict__ Vb, int64_t n, int m, const V *__restrict__ w,
                                              double *partial, const double *dh = nullptr,
                                              const double *dn = nullptr, int dm = 0) {
    __shared__ double sh[2 * kVB];
    if (dh && dgks_skip(dh, dm, dn)) return;
    gram_body<V>(Vb, n, m, w, partial, blockIdx.x, gridDim.x, sh);
}

__global__ __launch_bounds__(kVB) void k_reduce(const double *partial, int nb, double *out,
                                                const double *dh = nullptr, const double *dn = nullptr, int dm = 0,
                                                int copy_norm = 0) {
    __shared__ double sh[2 * kVB];
    if (dh && dgks_skip(dh, dm, dn)) {
        if (threadIdx.x == 0) {
            out[2 * blockIdx.x] = copy_norm ? dn[0] : 0.0;
            out[2 * blockIdx.x + 1] = 0.0;
        }
        return;
    }
    reduce_body(partial, nb, out, blockIdx.x, sh);
}

template <typename V>
__global__ __launch_bounds__(kVB) void k_update(const V *__restrict__ Vb, int64_t n, int m, const double *h,
                                                V *__restrict__ w, double *partial, const double *dh = nullptr,
                                                const double *dn = nullptr, int dm = 0) {
    __shared__ double sh[2 * kVB];
    if (dh && dgks_skip(dh, dm, dn)) return;
    update_body<V>(Vb, n, m, h, w, partial, blockIdx.x, gridDim.x, sh);
}

// The fused first-pass update and second-pass Gram product (m <= MX).
template <typename V, int MX>
__global__ __launch_bounds__(kVB) void k_update_gram(const V *__restrict__ Vb, int64_t n, int m, const double *h,
                                                     V *__restrict__ w, double *partial) {
    __shared__ double sh[2 * kVB];
    update_gram_body<V, MX>(Vb, n, m, h, w, partial, blockIdx.x, gridDim.x, sh);
}

// The second pass's update after k_update_gram: raw = the second pass's
// reduced Gram coefficients (m complex) followed by |w'|^2.  DGKS decides on
// the device: when no second pass is needed w stays as it is.  Block 0
// writes the second pass's coefficients (zeros when skipped) to h2 and
// |w'|^2 to n2a.
template <typename V>
__global__ __launch_bounds__(kVB) void k_update_dgks(const V *__restrict__ Vb, int64_t n, int m, const double *h1,
                                                     const double *raw, V *__restrict__ w, double *partial,
                                                     double *h2, double *n2a) {
    __shared__ double sh[2 * kVB];
    const bool skip = dgks_skip(h1, m, raw + 2 * m);
    if (blockIdx.x == 0) {
        for (int t = threadIdx.x; t < 2 * m; t += kVB) h2[t] = skip ? 0.0 : raw[t];
        if (threadIdx.x == 0) {
            n2a[0] = raw[2 * m];
            n2a[1] = 0.0;
        }
    }
    if (skip) return;
    update_body<V>(Vb, n, m, raw, w, partial, blockIdx.x, gridDim.x, sh);
}

// ------------------------------------------ small sectors, one workgroup --
// A whole run of Krylov steps of a small assembled sector in one launch of
// one workgroup (kSmallT threads; thread t owns rows t, t + kSmallT, ...):
// the same step sequence and device scalars as the per-operation launches
// (DGKS cycle of the thick-restart Lanczos, three-term Lanczos or two-pass
// Arnoldi of the expm factorization), with workgroup-wide sums in a fixed
// order instead of per-block partials.  Phases meet at workgroup barriers;
// every value a phase reads from global memory was written before the
// barrier that opens it.
constexpr int kSmallT = 512;
constexpr int kSmallW = kSmallT / 64;
constexpr int kSmallC = 8;    // columns per Gram pass

struct SmallKrylov {
    int mode;           // SMALL_DGKS, SMALL_LANCZOS, SMALL_ARNOLDI
    int64_t n;
    int steps, j0;      // DGKS: cycle length m and first step; else the factorization's steps
    char *V;            // basis, column-major with ld n
    const void *vin;    // factorization: starting vector
    double *dd;         // device scalars (layouts of the per-operation forms)
    int per;            // DGKS: scalars per step
    const uint64_t *sptr;
    const uint32_t *width, *cols;
    SellW wt;
};

// sum over the workgroup of every thread's (re, im): fixed order (wavefront
// butterflies, then the wavefront sums in order); every thread gets it
__device__ __forceinline__ void wg_sum2(double &re, double &im, double *red) {
    re = wave_sum(re);
    im = wave_sum(im);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[w] = re;
        red[kSmallW + w] = im;
    }
    __syncthreads();
    re = 0;
    im = 0;
    for (int i = 0; i < kSmallW; i++) {
        re += red[i];
        im += red[kSmallW + i];
    }
}

// h[c] = <V_c, w> for c < m (conj(V) . w), to LDS sh and global out
template <typename V>
__device__ void wg_gram(const V *Vb, int64_t n, int m, const V *w, double *sh, double *out, double *red) {
    using T = VTraits<V>;
    for (int c0 = 0; c0 < m; c0 += kSmallC) {
        double ar[kSmallC], ai[kSmallC];
#pragma unroll
        for (int c = 0; c < kSmallC; c++) ar[c] = ai[c] = 0.0;
        for (int64_t r = threadIdx.x; r < n; r += kSmallT) {
            double wr, wi;
            T::load(w, r, wr, wi);
#pragma unroll
            for (int c = 0; c < kSmallC; c++) {
                if (c0 + c < m) {
                    double vr, vi;
                    T::load(Vb, (int64_t)(c0 + c) * n + r, vr, vi);
                    ar[c] += vr * wr + vi * wi;
                    ai[c] += vr * wi - vi * wr;
                }
            }
        }
        // every column's wavefront butterflies at once, then the wavefront
        // sums of each column in order
#pragma unroll
        for (int c = 0; c < kSmallC; c++) {
            ar[c] = wave_sum(ar[c]);
            ai[c] = wave_sum(ai[c]);
        }
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
            for (int c = 0; c < kSmallC; c++) {
                red[(2 * c) * kSmallW + wv] = ar[c];
                red[(2 * c + 1) * kSmallW + wv] = ai[c];
            }
        }
        __syncthreads();
        const int t = threadIdx.x;
        if (t < 2 * kSmallC && c0 + t / 2 < m) {
            double v = 0;
            for (int i = 0; i < kSmallW; i++) v += red[t * kSmallW + i];
            sh[2 * c0 + t] = v;
            if (out) out[2 * c0 + t] = v;
        }
        __syncthreads();
    }
}

// w -= V h (h: m complex in LDS), then |w|^2 of the stored w
template <typename V>
__device__ double wg_update(const V *Vb, int64_t n, int m, const double *h, V *w, double *red) {
    using T = VTraits<V>;
    double nrm = 0;
    for (int64_t r = threadIdx.x; r < n; r += kSmallT) {
        double wr, wi;
        T::load(w, r, wr, wi);
        // columns in groups of kSmallC: the group's loads in flight together,
        // then the updates in column order
        for (int c0 = 0; c0 < m; c0 += kSmallC) {
            double vr[kSmallC], vi[kSmallC];
#pragma unroll
            for (int u = 0; u < kSmallC; u++) {
                vr[u] = vi[u] = 0.0;
                if (c0 + u < m) T::load(Vb, (int64_t)(c0 + u) * n + r, vr[u], vi[u]);
            }
#pragma unroll
            for (int u = 0; u < kSmallC; u++) {
                if (c0 + u < m) {
                    const double hr = h[2 * (c0 + u)], hi = h[2 * (c0 + u) + 1];
                    wr -= vr[u] * hr - vi[u] * hi;
                    wi -= vr[u] * hi + vi[u] * hr;
                }
            }
        }
        T::store(w, r, wr, wi);
        T::load(w, r, wr, wi);
        nrm += wr * wr + wi * wi;
    }
    double dummy = 0;
    wg_sum2(nrm, dummy, red);
    return nrm;
}

template <typename V>
__device__ double wg_norm2(const V *x, int64_t n, double *red) {
    using T = VTraits<V>;
    double nrm = 0;
    for (int64_t r = threadIdx.x; r < n; r += kSmallT) {
        double a, b;
        T::load(x, r, a, b);
        nrm += a * a + b * b;
    }
    double dummy = 0;
    wg_sum2(nrm, dummy, red);
    return nrm;
}

// y = a x + b y (b_zero: y = a x), real a, b
template <typename V>
__device__ void wg_axpby(double a, const V *x, double b, bool b_zero, V *y, int64_t n) {
    using T = VTraits<V>;
    for (int64_t r = threadIdx.x; r < n; r += kSmallT) {
        double xr, xi, yr = 0, yi = 0;
        T::load(x, r, xr, xi);
        if (!b_zero) T::load(y, r, yr, yi);
        const double orr = a * xr - 0.0 * xi + (b_zero ? 0.0 : b * yr - 0.0 * yi);
        const double oi = a * xi + 0.0 * xr + (b_zero ? 0.0 : b * yi + 0.0 * yr);
        T::store(y, r, orr, oi);
    }
}

template <typename V>
__device__ void wg_matvec(const SmallKrylov &K, const V *x, V *y) {
    using T = VTraits<V>;
    for (int64_t r = threadIdx.x; r < K.n; r += kSmallT) {
        double ar = 0, ai = 0;
        sell_row<V>(K.cols, K.wt, K.sptr[(uint64_t)r >> 6] + (r & 63), K.width[(uint64_t)r >> 6], x, 0, ar,
                    ai);
        T::store(y, r, ar, ai);
    }
}

template <typename V>
__global__ __launch_bounds__(kSmallT) void k_krylov_small(SmallKrylov K) {
    __shared__ double red[2 * kSmallC * kSmallW];
    __shared__ double sh1[2 * (kSmallMaxCols + 1)], sh2[2 * (kSmallMaxCols + 1)];
    const int64_t n = K.n;
    const size_t esz = sizeof(V);
    auto col = [&](int j) { return (V *)(K.V + (size_t)j * n * esz); };
    if (K.mode == SMALL_DGKS) {
        const int mc = K.steps;
        for (int j = K.j0; j < mc; j++) {
            double *h1 = K.dd + (size_t)K.per * j, *h2 = h1 + 2 * (mc + 1), *n2 = h2 + 2 * (mc + 1), *n2a = n2 + 2;
            V *w = col(j + 1);
            const int m = j + 1;
            wg_matvec<V>(K, col(j), w);
            __syncthreads();
            wg_gram<V>(col(0), n, m, w, sh1, h1, red);
            const double nw1 = wg_update<V>(col(0), n, m, sh1, w, red);
            __syncthreads();
            wg_gram<V>(col(0), n, m, w, sh2, nullptr, red);
            // ARPACK's DGKS test (dgks_skip): |w'|^2 > 0.717^2 (|w'|^2 + |h1|^2)
            double hh = 0;
            for (int c = 0; c < m; c++) hh += sh1[2 * c] * sh1[2 * c] + sh1[2 * c + 1] * sh1[2 * c + 1];
            const double e2 = 0.717 * 0.717;
            const bool skip = nw1 > e2 * (nw1 + hh);
            if (threadIdx.x == 0) {
                for (int c = 0; c < m; c++) {
                    h2[2 * c] = skip ? 0.0 : sh2[2 * c];
                    h2[2 * c + 1] = skip ? 0.0 : sh2[2 * c + 1];
                }
                n2a[0] = nw1;
                n2a[1] = 0.0;
            }
            double nw = nw1;
            if (!skip) {
                nw = wg_update<V>(col(0), n, m, sh2, w, red);
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                n2[0] = nw;
                n2[1] = 0.0;
            }
            wg_axpby<V>(1.0 / sqrt(nw), w, 0.0, true, w, n);
            __syncthreads();
        }
        return;
    }
    // the expm factorization (FactScalars layout: beta2 at 0, then per step
    // h1, h2 (2 (steps + 1) each) and |w|^2)
    const int steps = K.steps;
    const size_t per = (size_t)4 * (steps + 1) + 2;
    const double b2 = wg_norm2<V>((const V *)K.vin, n, red);
    if (threadIdx.x == 0) {
        K.dd[0] = b2;
        K.dd[1] = 0.0;
    }
    wg_axpby<V>(1.0 / sqrt(b2), (const V *)K.vin, 0.0, true, col(0), n);
    __syncthreads();
    double nprev = 0;
    for (int c = 0; c < steps; c++) {
        double *h1 = K.dd + 2 + (size_t)c * per, *h2 = h1 + 2 * (steps + 1), *n2 = h2 + 2 * (steps + 1);
        V *w = col(c + 1);
        wg_matvec<V>(K, col(c), w);
        __syncthreads();
        double nw;
        if (K.mode == SMALL_LANCZOS) {
            if (c > 0) {
                wg_axpby<V>(-sqrt(nprev), col(c - 1), 1.0, false, w, n);
                __syncthreads();
            }
            wg_gram<V>(col(c), n, 1, w, sh1, h1, red);
            nw = wg_update<V>(col(c), n, 1, sh1, w, red);
        } else {
            wg_gram<V>(col(0), n, c + 1, w, sh1, h1, red);
            wg_update<V>(col(0), n, c + 1, sh1, w, red);
            __syncthreads();
            wg_gram<V>(col(0), n, c + 1, w, sh2, h2, red);
            nw = wg_update<V>(col(0), n, c + 1, sh2, w, red);
        }
        if (threadIdx.x == 0) {
            n2[0] = nw;
            n2[1] = 0.0;
        }
        __syncthreads();
        if (c + 1 < steps) {
            wg_axpby<V>(1.0 / sqrt(nw), w, 0.0, true, w, n);
            __syncthreads();
        }
        nprev = nw;
    }
}

// ------------------------------------------------- block Gram on MFMA ----
// G = V^H W for V (n x m) and W (n x p), column-major with ld n: partial sums
// per row block on the matrix cores (v_mfma_f64_16x16x4f64).  One block owns
// a group of TM x TP output tiles of 16 x 16 (columns i0 .. i0 + 16 TM of V,
// j0 .. j0 + 16 TP of W, blockIdx.y numbering the groups), so each row of V
// and W is read once per group.  The block walks its row range in chunks of
// 64 R rows (R = 2 for FP32 storage, so a lane keeps as many bytes in flight
// as in FP64): the next chunk's values are loaded into registers (coalesced,
// 64 consecutive rows per column per load) while the matrix cores work on the
// current chunk from LDS, where FP32 values stay FP32 and are widened exactly
// as operands are read.  Wavefront w multiplies rows 16 R w .. 16 R w + 16 R - 1
// of a chunk as 4 R steps of K = 4 into every tile of the group.  Operand
// layout as in k_combine_mfma: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator register i of lane l is D[l / 16 + 4 i][l % 16].
// Complex: Re G = Vr'Wr + Vi'Wi, Im G = Vr'Wi - Vi'Wr.  The four wavefront
// tiles are summed in a fixed order; partial[(e nb + b) 2] holds output
// element e = i + m j, reduced over blocks by k_reduce.
template <typename V> struct GramStage {
    static constexpr bool f32 = std::is_same<V, float>::value || std::is_same<V, cplx_f>::value;
    using S = typename std::conditional<f32, float, double>::type;  // LDS element
    static constexpr int R = f32 ? 2 : 1;                            // rows per lane per chunk
    static constexpr int RC = 64 * R;                                // rows per chunk
    // LDS row stride: operand reads by (la, lk) land on distinct banks
    static constexpr int P = f32 ? RC + 4 : RC + 1;
};

template <typename V, int TM, int TP>
__global__ __launch_bounds__(kVB, 2) void k_gram_mfma(const V *__restrict__ Vb, int64_t n, int m,
                                                      const V *__restrict__ Wb, int p, int groups_m,
                                                      double *__restrict__ partial) {
    using T = VTraits<V>;
    using G = GramStage<V>;
    using S = typename G::S;
    constexpr bool CPLX = T::cplx;
    constexpr int NP = CPLX ? 2 : 1;             // planes: re (, im)
    constexpr int R = G::R, RC = G::RC, P = G::P;
    constexpr int CV = 16 * TM, CW = 16 * TP;    // columns staged per chunk
    constexpr int QV = CV / 4, QW = CW / 4;      // columns per lane
    constexpr size_t kStage = sizeof(S) * NP * (CV + CW) * P;
    constexpr size_t kRed = sizeof(double) * 2 * 4 * 256;
    __shared__ __align__(16) unsigned char smem[kStage > kRed ? kStage : kRed];
    S *sv = (S *)smem, *sw = sv + NP * CV * P;
    const int b = blockIdx.x, nb = gridDim.x;
    const int gi = blockIdx.y % groups_m, gj = blockIdx.y / groups_m;
    const int i0 = CV * gi, j0 = CW * gj;
    const int mc = m - i0 < CV ? m - i0 : CV, pc = p - j0 < CW ? p - j0 : CW;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rr = threadIdx.x & 63, cq = threadIdx.x >> 6;  // staging row / first column
    int64_t r0, r1;
    row_range(n, b, nb, r0, r1);
    S pv[NP][R][QV], pw[NP][R][QW];
    // lane t stages rows t % 64 + 64 u of columns t / 64 + 4 q; absent entries are zero
    auto fetch = [&](int64_t c0) {
#pragma unroll
        for (int u = 0; u < R; u++) {
            const int64_t r = c0 + rr + 64 * u;
            const bool in = r < r1;
#pragma unroll
            for (int q = 0; q < QV; q++) {
                const int c = cq + 4 * q;
                double re = 0, im = 0;
                if (in && c < mc) T::load(Vb, (int64_t)(i0 + c) * n + r, re, im);
                pv[0][u][q] = (S)re;
                if (CPLX) pv[NP - 1][u][q] = (S)im;
            }
#pragma unroll
            for (int q = 0; q < QW; q++) {
                const int c = cq + 4 * q;
                double re = 0, im = 0;
                if (in && c < pc) T::load(Wb, (int64_t)(j0 + c) * n + r, re, im);
                pw[0][u][q] = (S)re;
                if (CPLX) pw[NP - 1][u][q] = (S)im;
            }
        }
    };
    v4d acc_r[TM][TP], acc_i[TM][TP];
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
        for (int c = 0; c < TP; c++) {
            acc_r[a][c] = v4d{0, 0, 0, 0};
            acc_i[a][c] = v4d{0, 0, 0, 0};
        }
    const int la = lane & 15, lk = lane >> 4;
    if (r0 < r1) fetch(r0);
    for (int64_t c0 = r0; c0 < r1; c0 += RC) {
        __syncthreads();  // every wavefront is done with the previous chunk
#pragma unroll
        for (int h = 0; h < NP; h++)
#pragma unroll
            for (int u = 0; u < R; u++) {
#pragma unroll
                for (int q = 0; q < QV; q++) sv[(h * CV + cq + 4 * q) * P + rr + 64 * u] = pv[h][u][q];
#pragma unroll
                for (int q = 0; q < QW; q++) sw[(h * CW + cq + 4 * q) * P + rr + 64 * u] = pw[h][u][q];
            }
        __syncthreads();
        if (c0 + RC < r1) fetch(c0 + RC);  // in flight during the products below
#pragma unroll
        for (int st = 0; st < 4 * R; st++) {
            const int kr = 16 * R * wv + 4 * st + lk;
            double ar[TM], ai[TM], br[TP], bi[TP];
#pragma unroll
            for (int a = 0; a < TM; a++) {
                ar[a] = (double)sv[(16 * a + la) * P + kr];
                ai[a] = CPLX ? (double)sv[(CV + 16 * a + la) * P + kr] : 0.0;
            }
#pragma unroll
            for (int c = 0; c < TP; c++) {
                br[c] = (double)sw[(16 * c + la) * P + kr];
                bi[c] = CPLX ? (double)sw[(CW + 16 * c + la) * P + kr] : 0.0;
            }
#pragma unroll
            for (int a = 0; a < TM; a++)
#pragma unroll
                for (int c = 0; c < TP; c++) {
                    acc_r[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[c], acc_r[a][c], 0, 0, 0);
                    if (CPLX) {
                        acc_r[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bi[c], acc_r[a][c], 0, 0, 0);
                        acc_i[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[c], acc_i[a][c], 0, 0, 0);
                        acc_i[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], br[c], acc_i[a][c], 0, 0, 0);
                    }
                }
        }
    }
    // per tile: D[l / 16 + 4 q][l % 16] of each wavefront -> LDS, then a
    // fixed-order sum over the four wavefronts
    double *sred = (double *)smem;  // [wavefront][re, im][256]
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
        for (int c = 0; c < TP; c++) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int e = (lk + 4 * q) * 16 + la;
                sred[(wv * 2) * 256 + e] = acc_r[a][c][q];
                sred[(wv * 2 + 1) * 256 + e] = acc_i[a][c][q];
            }
            __syncthreads();
            const int e = threadIdx.x, ii = 16 * a + (e >> 4), jj = 16 * c + (e & 15);
            if (ii < mc && jj < pc) {
                const double gr = (sred[0 * 256 + e] + sred[2 * 256 + e]) + (sred[4 * 256 + e] + sred[6 * 256 + e]);
                const double gim = (sred[1 * 256 + e] + sred[3 * 256 + e]) + (sred[5 * 256 + e] + sred[7 * 256 + e]);
                const int64_t g = (int64_t)(i0 + ii) + (int64_t)m * (j0 + jj);
                partial[(g * nb + b) * 2] = gr;
                partial[(g * nb + b) * 2 + 1] = CPLX ? gim : 0.0;
            }
        }
}

template <typename V>
__global__ __launch_bounds__(kVB) void k_norm2(const V *__restrict__ x, int64_t n, double *partial) {
    __shared__ double sh[2 * kVB];
    norm2_body<V>(x, n, partial, blockIdx.x, gridDim.x, sh);
}

template <typename V>
__global__ void k_axpby(double ar, double ai, const V *__restrict__ x, double br, double bi, V *__restrict__ y,
                        int64_t n, int b_zero, const double *acoef, int cmode) {
    axpby_body<V>(ar, ai, x, br, bi, y, n, b_zero, blockIdx.x, gridDim.x, acoef, cmode);
}

template <typename V>
__global__ void k_combine(const V *__restrict__ Vb, int64_t n, int m, const double *C, int k, V *__restrict__ Y) {
    combine_body<V>(Vb, n, m, C, k, Y, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------ batched launches --
// Block -> (job, local block) by the jobs' first-block offsets (ascending).
__device__ __forceinline__ int find_job(const VecJobs &J, int &b) {
    int lo = 0, hi = J.njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)blockIdx.x >= J.job[mid].b0) lo = mid;
        else hi = mid - 1;
    }
    b = (int)blockIdx.x - J.job[lo].b0;
    return lo;
}

template <typename V>
__device__ __forceinline__ void vec_job_block(const VecJob &q, int b, int op, double *sh) {
    switch (op) {
    case VOP_GRAM: gram_body<V>((const V *)q.V, q.n, q.m, (const V *)q.x, q.partial, b, q.nb, sh); break;
    case VOP_REDUCE: reduce_body(q.partial, q.nb, q.aux, b, sh); break;
    case VOP_UPDATE: update_body<V>((const V *)q.V, q.n, q.m, q.aux, (V *)q.w, q.partial, b, q.nb, sh); break;
    case VOP_NORM2: norm2_body<V>((const V *)q.x, q.n, q.partial, b, q.nb, sh); break;
    case VOP_AXPBY:
        axpby_body<V>(q.ar, q.ai, (const V *)q.x, q.br, q.bi, (V *)q.w, q.n, q.b_zero, b, q.nb, q.acoef, q.cmode);
        break;
    case VOP_COMBINE: combine_body<V>((const V *)q.V, q.n, q.m, q.aux, q.k, (V *)q.w, b, q.nb); break;
    }
}

template <typename V>
__global__ __launch_bounds__(kVB) void kb_vec(VecJobs J, int op) {
    __shared__ double sh[2 * kVB];
    int b;
    const VecJob &q = J.job[find_job(J, b)];
    vec_job_block<V>(q, b, op, sh);
}

// The fused first-pass update and second-pass Gram product (update_gram_body)
// over a device job table, every job with the same column count m <= MX:
// job j owns blocks [j nbmax, (j+1) nbmax) and uses the first tab[j].b0.
template <typename V, int MX>
__global__ __launch_bounds__(kVB) void kb_updgram_tab(const VecJob *__restrict__ tab, int nbmax) {
    __shared__ double sh[2 * kVB];
    const int j = (int)blockIdx.x / nbmax, b = (int)blockIdx.x - j * nbmax;
    const VecJob q = tab[j];
    if (b >= q.b0) return;
    update_gram_body<V, MX>((const V *)q.V, q.n, q.m, q.aux, (V *)q.w, q.partial, b, q.nb, sh);
}

// The same over a device job table: job j owns blocks [j nbmax, (j+1) nbmax)
// and uses the first tab[j].b0 of them (b0 holds its block count here).
template <typename V>
__global__ __launch_bounds__(kVB) void kb_vec_tab(const VecJob *__restrict__ tab, int nbmax, int op) {
    __shared__ double sh[2 * kVB];
    const int j = (int)blockIdx.x / nbmax, b = (int)blockIdx.x - j * nbmax;
    const VecJob q = tab[j];
    if (b >= q.b0) return;
    vec_job_block<V>(q, b, op, sh);
}

template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ x, B *__restrict__ y, int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double re, im;
        VTraits<A>::load(x, r, re, im);
        VTraits<B>::store(y, r, re, im);
    }
}

// uniform(-1, 1) entries from a counter-based hash (splitmix64) of (seed, index)
__device__ __forceinline__ double hash_uniform(uint64_t seed, uint64_t i) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

template <typename V>
__global__ void k_random(V *__restrict__ y, int64_t n, uint64_t seed) {
    using T = VTraits<V>;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double re = hash_uniform(seed, 2 * (uint64_t)r);
        double im = T::cplx ? hash_uniform(seed, 2 * (uint64_t)r + 1) : 0.0;
        T::store(y, r, re, im);
    }
}

int ew_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }


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

void ensure(VecWork &wk, size_t partial, size_t result) {
    if (wk.partial.n < partial) wk.partial.alloc(std::max(partial, (size_t)2 * kMaxBlocks * 64), wk.s);
    if (wk.result.n < result) wk.result.alloc(std::max(result, (size_t)2 * 512), wk.s);
}

}  // namespace

int ew_blocks(int64_t n) { return ew_grid(n); }

bool combine_on_mfma(int dtype, const void *Vmat, int64_t n, int m, const double *C, int k, void *Y,
                     hipStream_t s) {
    // many output columns (a Krylov basis against a time grid or a block of
    // Ritz vectors): a GEMM for the matrix cores; FP64 storage only
    if (k < kMfmaMinCols || n < 16 || (dtype != QS_F64 && dtype != QS_C128)) return false;
    const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((k + kMfmaCols - 1) / kMfmaCols));
    if (dtype == QS_C128)
        hipLaunchKernelGGL((k_combine_mfma<true>), grid, dim3(256), 0, s, (const double *)Vmat, n, m, C, k,
                           (double *)Y);
    else
        hipLaunchKernelGGL((k_combine_mfma<false>), grid, dim3(256), 0, s, (const double *)Vmat, n, m, C, k,
                           (double *)Y);
    QS_HIP(hipGetLastError());
    return true;
}

void launch_vec_jobs(int dtype, int op, std::vector<VecJob> &jobs, const std::vector<int> &blocks, hipStream_t s,
                     JobTable *table) {
    if (table && jobs.size() > (size_t)kMaxVecJobs) {
        int nbmax = 0;
        for (int nb : blocks) nbmax = std::max(nbmax, nb);
        if (nbmax == 0) return;
        if ((int64_t)nbmax * (int64_t)jobs.size() > INT32_MAX) fail(QS_EARG, "internal: batched grid too large");
        for (size_t i = 0; i < jobs.size(); i++) jobs[i].b0 = blocks[i];
        const VecJob *tab = (const VecJob *)table->stage(jobs.data(), jobs.size() * sizeof(VecJob), s);
        const int grid = nbmax * (int)jobs.size();
        by_dtype(dtype, [&](auto *tag) {
            using V = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL((kb_vec_tab<V>), dim3(grid), dim3(kVB), 0, s, tab, nbmax, op);
        });
        QS_HIP(hipGetLastError());
        return;
    }
    for (size_t first = 0; first < jobs.size(); first += kMaxVecJobs) {
        VecJobs J;
        J.njobs = (int)std::min<size_t>(kMaxVecJobs, jobs.size() - first);
        int total = 0;
        for (int i = 0; i < J.njobs; i++) {
            J.job[i] = jobs[first + i];
            J.job[i].b0 = total;
            total += blocks[first + i];
        }
        if (total == 0) continue;
        by_dtype(dtype, [&](auto *tag) {
            using V = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL((kb_vec<V>), dim3(total), dim3(kVB), 0, s, J, op);
        });
        QS_HIP(hipGetLastError());
    }
}

bool launch_updgram_jobs(int dtype, std::vector<VecJob> &jobs, hipStream_t s, JobTable *table) {
    if (jobs.empty()) return true;
    const int m = jobs[0].m;
    if (m > kDgksFuse) return false;
    int nbmax = 0;
    for (auto &j : jobs) {
        if (j.m != m) fail(QS_EARG, "internal: fused Gram jobs of different column counts");
        nbmax = std::max(nbmax, j.nb);
        j.b0 = j.nb;
    }
    if ((int64_t)nbmax * (int64_t)jobs.size() > INT32_MAX) fail(QS_EARG, "internal: batched grid too large");
    const VecJob *tab = (const VecJob *)table->stage(jobs.data(), jobs.size() * sizeof(VecJob), s);
    const int grid = nbmax * (int)jobs.size();
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        if (m <= 8) hipLaunchKernelGGL((kb_updgram_tab<V, 8>), dim3(grid), dim3(kVB), 0, s, tab, nbmax);
        else if (m <= 16) hipLaunchKernelGGL((kb_updgram_tab<V, 16>), dim3(grid), dim3(kVB), 0, s, tab, nbmax);
        else hipLaunchKernelGGL((kb_updgram_tab<V, kDgksFuse>), dim3(grid), dim3(kVB), 0, s, tab, nbmax);
    });
    QS_HIP(hipGetLastError());
    return true;
}

int vec_blocks(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (n + 4 * kVB - 1) / (4 * kVB)));
}

void gram_dots(int dtype, const void *Vmat, int64_t n, int m, const void *w, VecWork &wk, hipStream_t s,
               std::vector<cd> &h) {
    int nb = vec_blocks(n);
    ensure(wk, (size_t)2 * nb * m, (size_t)2 * m);
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_gram<V>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, (const V *)w,
                           wk.partial.get());
    });
    hipLaunchKernelGGL(k_reduce, dim3(m), dim3(kVB), 0, s, wk.partial.get(), nb, wk.result.get());
    QS_HIP(hipGetLastError());
    wk.host.resize((size_t)2 * m);
    QS_HIP(hipMemcpyAsync(wk.host.data(), wk.result.get(), sizeof(double) * 2 * m, hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    h.resize(m);
    for (int j = 0; j < m; j++) h[j] = cd(wk.host[2 * j], wk.host[2 * j + 1]);
}

double gram_update(int dtype, const void *Vmat, int64_t n, int m, const std::vector<cd> &h, void *w, VecWork &wk,
                   hipStream_t s) {
    int nb = vec_blocks(n);
    ensure(wk, (size_t)2 * nb + 2 * m + 2, (size_t)2 * m + 2);
    // h goes after the partials in the same buffer
    std::vector<double> hh((size_t)2 * m);
    for (int j = 0; j < m; j++) {
        hh[2 * j] = h[j].real();
        hh[2 * j + 1] = h[j].imag();
    }
    double *dh = wk.partial.get() + 2 * nb;
    if (m > 0) QS_HIP(hipMemcpyAsync(dh, hh.data(), sizeof(double) * 2 * m, hipMemcpyHostToDevice, s));
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_update<V>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, dh, (V *)w,
                           wk.partial.get());
    });
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kVB), 0, s, wk.partial.get(), nb, wk.result.get());
    QS_HIP(hipGetLastError());
    double out[2];
    QS_HIP(hipMemcpyAsync(out, wk.result.get(), sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    return out[0];
}

double vec_norm2(int dtype, const void *x, int64_t n, VecWork &wk, hipStream_t s) {
    int nb = vec_blocks(n);
    ensure(wk, (size_t)2 * nb, 2);
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_norm2<V>), dim3(nb), dim3(kVB), 0, s, (const V *)x, n, wk.partial.get());
    });
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kVB), 0, s, wk.partial.get(), nb, wk.result.get());
    QS_HIP(hipGetLastError());
    double out[2];
    QS_HIP(hipMemcpyAsync(out, wk.result.get(), sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    return out[0];
}

void vec_scale_copy(int dtype, const void *x, void *y, int64_t n, cd scale, hipStream_t s) {
    vec_axpby(dtype, scale, x, cd(0, 0), y, n, s);
}

void vec_axpby(int dtype, cd a, const void *x, cd b, void *y, int64_t n, hipStream_t s) {
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_axpby<V>), dim3(ew_grid(n)), dim3(256), 0, s, a.real(), a.imag(), (const V *)x,
                           b.real(), b.imag(), (V *)y, n, (int)(b == cd(0, 0)), (const double *)nullptr, (int)CO_HOST);
    });
    QS_HIP(hipGetLastError());
}

// ------------------------------------------- device-scalar variants -------
// The same kernels with every scalar left in device memory: no host round
// trip, so a whole Krylov factorization is queued on the stream at once.

void vec_reserve(VecWork &wk, int64_t n, int mmax) {
    const int nb = vec_blocks(n);
    ensure(wk, (size_t)2 * nb * std::max(1, mmax) + 2 * (size_t)std::max(1, mmax) + 2,
           (size_t)2 * std::max(1, mmax) + 2);
}

void gram_dots_dev(int dtype, const void *Vmat, int64_t n, int m, const void *w, VecWork &wk, double *out,
                   hipStream_t s, const double *dgks_h1, const double *dgks_n2) {
    const int nb = vec_blocks(n);
    if (wk.partial.n < (size_t)2 * nb * m) fail(QS_EARG, "internal: vector workspace not reserved");
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_gram<V>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, (const V *)w,
                           wk.partial.get(), dgks_h1, dgks_n2, m);
    });
    hipLaunchKernelGGL(k_reduce, dim3(m), dim3(kVB), 0, s, wk.partial.get(), nb, out, dgks_h1, dgks_n2, m, 0);
    QS_HIP(hipGetLastError());
}

void vec_reduce_dev(const double *partial, int nb, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kVB), 0, s, partial, nb, out, nullptr, nullptr, 0, 0);
    QS_HIP(hipGetLastError());
}

void gram_update_dev(int dtype, const void *Vmat, int64_t n, int m, const double *h, void *w, VecWork &wk,
                     double *nrm2_out, hipStream_t s, const double *dgks_h1, const double *dgks_n2) {
    const int nb = vec_blocks(n);
    if (wk.partial.n < (size_t)2 * nb) fail(QS_EARG, "internal: vector workspace not reserved");
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_update<V>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, h, (V *)w,
                           wk.partial.get(), dgks_h1, dgks_n2, m);
    });
    if (nrm2_out)
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kVB), 0, s, wk.partial.get(), nb, nrm2_out, dgks_h1, dgks_n2, m,
                           1);
    QS_HIP(hipGetLastError());
}

void gram_dgks_dev(int dtype, const void *Vmat, int64_t n, int m, void *w, VecWork &wk, double *h1, double *h2,
                   double *n2a, double *n2, hipStream_t s) {
    const int nb = vec_blocks(n);
    gram_dots_dev(dtype, Vmat, n, m, w, wk, h1, s);
    const size_t need = (size_t)2 * nb * (m + 1) + (size_t)2 * (m + 1);
    if (m > kDgksFuse || wk.partial.n < need) {
        gram_update_dev(dtype, Vmat, n, m, h1, w, wk, n2a, s);
        gram_dots_dev(dtype, Vmat, n, m, w, wk, h2, s, h1, n2a);
        gram_update_dev(dtype, Vmat, n, m, h2, w, wk, n2, s, h1, n2a);
        return;
    }
    double *raw = wk.partial.get() + (size_t)2 * nb * (m + 1);
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        if (m <= 8)
            hipLaunchKernelGGL((k_update_gram<V, 8>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, h1, (V *)w,
                               wk.partial.get());
        else if (m <= 16)
            hipLaunchKernelGGL((k_update_gram<V, 16>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, h1, (V *)w,
                               wk.partial.get());
        else
            hipLaunchKernelGGL((k_update_gram<V, kDgksFuse>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, h1,
                               (V *)w, wk.partial.get());
    });
    hipLaunchKernelGGL(k_reduce, dim3(m + 1), dim3(kVB), 0, s, wk.partial.get(), nb, raw, (const double *)nullptr,
                       (const double *)nullptr, 0, 0);
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_update_dgks<V>), dim3(nb), dim3(kVB), 0, s, (const V *)Vmat, n, m, h1, raw, (V *)w,
                           wk.partial.get(), h2, n2a);
    });
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kVB), 0, s, wk.partial.get(), nb, n2, h1, raw + 2 * m, m, 1);
    QS_HIP(hipGetLastError());
}

void gram_block_dev(int dtype, const void *Vmat, int64_t n, int m, const void *Wmat, int p, VecWork &wk,
                    double *out, hipStream_t s) {
    if (m < 1 || p < 1) fail(QS_EARG, "internal: empty block Gram product");
    // tile groups of up to 2 x 2 tiles (32 x 32 outputs) per block
    const int tm = (m + 15) / 16, tp = (p + 15) / 16;
    const int TM = tm > 1 ? 2 : 1, TP = tp > 1 ? 2 : 1;
    const int gm = (tm + TM - 1) / TM, gp = (tp + TP - 1) / TP;
    int dev = 0, cus = 0;
    QS_HIP(hipGetDevice(&dev));
    QS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int nb = 0;
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        const void *kern = TM == 2 && TP == 2 ? (const void *)k_gram_mfma<V, 2, 2>
                           : TM == 2          ? (const void *)k_gram_mfma<V, 2, 1>
                           : TP == 2          ? (const void *)k_gram_mfma<V, 1, 2>
                                              : (const void *)k_gram_mfma<V, 1, 1>;
        // one round of resident blocks over the groups: no partly filled
        // last round; fewer blocks when n is small
        int occ = 0;
        QS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kVB, 0));
        const int per_group = std::max(1, occ * cus / (gm * gp));
        nb = std::max(1, std::min(vec_blocks(n), per_group));
        ensure(wk, (size_t)2 * nb * m * p, 2);
        void *args[] = {(void *)&Vmat, (void *)&n, (void *)&m, (void *)&Wmat, (void *)&p, (void *)&gm,
                        (void *)&wk.partial.p};
        QS_HIP(hipLaunchKernel(kern, dim3(nb, gm * gp), dim3(kVB), args, 0, s));
    });
    hipLaunchKernelGGL(k_reduce, dim3(m * p), dim3(kVB), 0, s, wk.partial.get(), nb, out, (const double *)nullptr,
                       (const double *)nullptr, 0, 0);
    QS_HIP(hipGetLastError());
}

void gram_block(int dtype, const void *Vmat, int64_t n, int m, const void *Wmat, int p, VecWork &wk, hipStream_t s,
                std::vector<cd> &G) {
    SBuf<double> d((size_t)2 * m * p, s);
    gram_block_dev(dtype, Vmat, n, m, Wmat, p, wk, d.get(), s);
    std::vector<double> h((size_t)2 * m * p);
    QS_HIP(hipMemcpyAsync(h.data(), d.get(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    QS_HIP(hipStreamSynchronize(s));
    G.resize((size_t)m * p);
    for (size_t e = 0; e < G.size(); e++) G[e] = cd(h[2 * e], h[2 * e + 1]);
}

void vec_norm2_dev(int dtype, const void *x, int64_t n, VecWork &wk, double *out, hipStream_t s) {
    const int nb = vec_blocks(n);
    if (wk.partial.n < (size_t)2 * nb) fail(QS_EARG, "internal: vector workspace not reserved");
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_norm2<V>), dim3(nb), dim3(kVB), 0, s, (const V *)x, n, wk.partial.get());
    });
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kVB), 0, s, wk.partial.get(), nb, out);
    QS_HIP(hipGetLastError());
}

void vec_axpby_dev(int dtype, const double *acoef, int cmode, const void *x, cd b, void *y, int64_t n,
                   hipStream_t s) {
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_axpby<V>), dim3(ew_grid(n)), dim3(256), 0, s, 0.0, 0.0, (const V *)x, b.real(),
                           b.imag(), (V *)y, n, (int)(b == cd(0, 0)), acoef, cmode);
    });
    QS_HIP(hipGetLastError());
}

void basis_combine(int dtype, const void *Vmat, int64_t n, int m, const std::vector<cd> &C, int k, void *Y,
                   hipStream_t s) {
    std::vector<double> cc((size_t)2 * m * k);
    for (size_t i = 0; i < (size_t)m * k; i++) {
        cc[2 * i] = C[i].real();
        cc[2 * i + 1] = C[i].imag();
    }
    SBuf<double> dc(cc.size(), s);
    QS_HIP(hipMemcpyAsync(dc.get(), cc.data(), sizeof(double) * cc.size(), hipMemcpyHostToDevice, s));
    if (!combine_on_mfma(dtype, Vmat, n, m, dc.get(), k, Y, s)) {
        by_dtype(dtype, [&](auto *tag) {
            using V = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL((k_combine<V>), dim3(ew_grid(n)), dim3(256), 0, s, (const V *)Vmat, n, m, dc.get(),
                               k, (V *)Y);
        });
    }
    QS_HIP(hipGetLastError());
    QS_HIP(hipStreamSynchronize(s));
}

cd rayleigh_numerator(OpHost &op, int dtype, const void *x, void *y, VecWork &wk) {
    const int g = matvec_grid(op.basis->Ns);
    ensure(wk, (size_t)2 * g, 2);
    launch_matvec(op, dtype, x, y, 1.0, 0.0, 0.0, 0.0, false, false, false, op_stream(op), wk.partial.get());
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kVB), 0, op_stream(op), wk.partial.get(), g, wk.result.get());
    QS_HIP(hipGetLastError());
    double out[2];
    QS_HIP(hipMemcpyAsync(out, wk.result.get(), sizeof(double) * 2, hipMemcpyDeviceToHost, op_stream(op)));
    QS_HIP(hipStreamSynchronize(op_stream(op)));
    return cd(out[0], out[1]);
}

void vec_random(int dtype, void *y, int64_t n, uint64_t seed, hipStream_t s) {
    by_dtype(dtype, [&](auto *tag) {
        using V = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_random<V>), dim3(ew_grid(n)), dim3(256), 0, s, (V *)y, n, seed);
    });
    QS_HIP(hipGetLastError());
}

void vec_convert(int from, const void *x, int to, void *y, int64_t n, hipStream_t s) {
    by_dtype(from, [&](auto *ta) {
        using A = std::remove_pointer_t<decltype(ta)>;
        by_dtype(to, [&](auto *tb) {
            using B = std::remove_pointer_t<decltype(tb)>;
            hipLaunchKernelGGL((k_convert<A, B>), dim3(ew_grid(n)), dim3(256), 0, s, (const A *)x, (B *)y, n);
        });
    });
    QS_HIP(hipGetLastError());
}

bool krylov_small_ok(const OpHost &op, int cols) {
    const BasisHost &b = *op.basis;
    return op.sell && !b.wide && b.Ns <= kSmallMaxN && cols <= kSmallMaxCols &&
           (op.lanes_per_row == 0 || op.lanes_per_row == QS_VARIANT_SELL);
}

void krylov_small(const OpHost &op, int dtype, int mode, int steps, int j0, void *V, const void *vin, double *dd,
                  int per, hipStream_t s) {
    const SellHost &t = *op.sell;
    if (t.cplx && !dtype_complex(dtype)) fail(QS_EARG, "a complex sector needs a complex vector dtype");
    SmallKrylov K;
    K.mode = mode;
    K.n = op.basis->Ns;
    K.steps = steps;
    K.j0 = j0;
    K.V = (char *)V;
    K.vin = vin;
    K.dd = dd;
    K.per = per;
    K.sptr = t.sptr.get();
    K.width = t.width.get();
    K.cols = t.cols.get();
    K.wt = t.weights();
    by_dtype(dtype, [&](auto *tag) {
        using VT = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_krylov_small<VT>), dim3(1), dim3(kSmallT), 0, s, K);
    });
    QS_HIP(hipGetLastError());
}

}  // namespace qs
