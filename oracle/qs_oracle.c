/*
 * qs_oracle.c -- CPU oracle for the spin-1/2 symmetry-sector Hamiltonian action.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * algorithms QuSpin.jl runs on its CPU path for spin-1/2 chains.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker or as the timed CPU baseline.  The product path
 * (libqsgpu.so) never links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   qo_fixed_weight_states  basis.jl:229-250      Gosper enumeration, ascending
 *   qo_sector_build         basis.jl:622-777      stage order k, p, z, pz, zA, zB
 *   qo_sector_build_general basis.jl:282-338,530-620  SpinBasisGeneral site maps, stages in block order
 *   qo_wsector_build        basis.jl:361-445,530-620  the same for up to 256 sites (UInt256 states)
 *                           symmetry_basis.jl:38-55   signed permutation S
 *                           symmetry_basis.jl:57-108  cyclic projector (orbit walk)
 *                           symmetry_basis.jl:110-213 eigenspace intersection P'SP
 *                           symmetry_basis.jl:215-232 representative = first max |P|
 *                           discrete_basis.jl:2386-2413 normalization = 1/|amp|^2
 *   qo_op_*                 operators.jl:214-263  term consolidation (first-seen order)
 *                           operators.jl:343-379  _apply_local
 *                           operators.jl:2372-2429 diagonal extraction (check_symm)
 *                           operators.jl:2554-2636 lift / parent apply / project
 *                           operators.jl:3097-3179 mul!(y, qlo, x, alpha, beta)
 *   qo_csr_*                operators.jl:381-576  triplet assembly, zero diagonals dropped
 *                           sparse_formats.jl:29-50,110-146 CSR storage and SpMV
 *
 * Floating-point emulation.  The reference decides representative labels by
 * comparing |P| entries that differ only by rounding.  The sector builder
 * therefore repeats the reference's operation sequence in IEEE double: Julia's
 * complex multiply (no fused multiply-add), Julia's robust complex division
 * (Baudin & Smith 2012), LinearAlgebra.generic_norm2 for orbit norms, the
 * Gustavson column-by-column accumulation order of SparseArrays' spmatmul, and
 * a correctly rounded hypot for |z|.  cis() uses the host libm cos/sin.
 * Orbit norms for orbits of 32 or more states go through BLAS nrm2 in the
 * reference; here they use the same sequential sum (parity unpinned for such
 * orbits, which occur only for L >= 32 with translations).
 *
 * Compile with -ffp-contract=off so that no expression is fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static cplx c_make(double re, double im) { cplx z; z.re = re; z.im = im; return z; }
static cplx c_mul(cplx a, cplx b)
{
    return c_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static cplx c_add(cplx a, cplx b) { return c_make(a.re + b.re, a.im + b.im); }
static cplx c_sub(cplx a, cplx b) { return c_make(a.re - b.re, a.im - b.im); }
static cplx c_conj(cplx a) { return c_make(a.re, -a.im); }

/* Julia Base: robust complex division for ComplexF64 (Baudin & Smith 2012). */
static double cdiv2(double a, double b, double c, double d, double r, double t)
{
    if (r != 0) {
        double br = b * r;
        return (br != 0) ? (a + br) * t : a * t + (b * t) * r;
    }
    return (a + d * (b / c)) * t;
}
static void cdiv1(double a, double b, double c, double d, double *p, double *q)
{
    double r = d / c;
    double t = 1.0 / (c + d * r);
    *p = cdiv2(a, b, c, d, r, t);
    *q = cdiv2(b, -a, c, d, r, t);
}
static cplx c_div(cplx z, cplx w)
{
    double a = z.re, b = z.im, c = w.re, d = w.im;
    double absa = fabs(a), absb = fabs(b);
    double ab = absa >= absb ? absa : absb;
    double absc = fabs(c), absd = fabs(d);
    double cd = absc >= absd ? absc : absd;
    const double eps = 2.220446049250313e-16;
    double halfov = 0.5 * 1.7976931348623157e308;
    double twouneps = 2.2250738585072014e-308 * 2.0 / eps;
    double bs = 2.0 / (eps * eps);
    double s = 1.0;
    if (ab >= halfov) { a *= 0.5; b *= 0.5; s *= 2.0; }
    if (cd >= halfov) { c *= 0.5; d *= 0.5; s *= 0.5; }
    if (ab <= twouneps) { a *= bs; b *= bs; s /= bs; }
    if (cd <= twouneps) { c *= bs; d *= bs; s *= bs; }
    double p, q;
    if (fabs(d) <= fabs(c)) {
        cdiv1(a, b, c, d, &p, &q);
    } else {
        cdiv1(b, a, d, c, &p, &q);
        q = -q;
    }
    return c_make(p * s, q * s);
}

/* Correctly rounded hypot, following Julia Base.Math.hypot on FMA hardware. */
static double jl_hypot(double x, double y)
{
    if (isinf(x) || isinf(y)) return INFINITY;
    double ax = fabs(x), ay = fabs(y);
    if (ay > ax) { double t = ax; ax = ay; ay = t; }
    const double eps = 2.220446049250313e-16;
    if (ay <= ax * sqrt(eps / 2)) return ax;
    /* eps(sqrt(floatmin(Float64))) = ulp(2^-511) = 2^-563 */
    double scale = ldexp(1.0, -563);
    if (ax > sqrt(1.7976931348623157e308 / 2)) {
        ax *= scale; ay *= scale; scale = 1.0 / scale;
    } else if (ay < sqrt(2.2250738585072014e-308)) {
        ax /= scale; ay /= scale;
    } else {
        scale = 1.0;
    }
    double h = sqrt(fma(ax, ax, ay * ay));
    double hsq = h * h, axsq = ax * ax;
    h -= (fma(-ay, ay, hsq - axsq) + fma(h, h, -hsq) - fma(ax, ax, -axsq)) / (2 * h);
    return h * scale;
}
static double c_abs(cplx z) { return jl_hypot(z.re, z.im); }
static double c_abs2(cplx z) { return z.re * z.re + z.im * z.im; }

/* Base.isapprox(x, y; atol, rtol) for complex numbers. */
static int c_isapprox(cplx x, cplx y, double atol, double rtol)
{
    if (x.re == y.re && x.im == y.im) return 1;
    double d = c_abs(c_sub(x, y));
    double m = c_abs(x);
    double my = c_abs(y);
    if (my > m) m = my;
    double tol = rtol * m;
    if (atol > tol) tol = atol;
    return d <= tol;
}

/* LinearAlgebra.generic_norm2 on a short complex vector (no rescaling branch
 * is taken for the unit-modulus coefficients that occur here). */
static double generic_norm2(const cplx *v, int n)
{
    double maxabs = 0;
    for (int i = 0; i < n; i++) {
        double a = c_abs(v[i]);
        if (a > maxabs) maxabs = a;
    }
    if (maxabs == 0 || isinf(maxabs)) return maxabs;
    double sum = c_abs2(v[0]);
    for (int i = 1; i < n; i++) sum += c_abs2(v[i]);
    return sqrt(sum);
}

/* Test hooks: the floating-point primitives above, exported so the tests
 * can pin them against exact rational arithmetic. */
double qo_prim_hypot(double x, double y) { return jl_hypot(x, y); }
void qo_prim_cdiv(double a, double b, double c, double d, double *out)
{
    cplx q = c_div(c_make(a, b), c_make(c, d));
    out[0] = q.re;
    out[1] = q.im;
}
double qo_prim_norm2(const double *v, int n) { return generic_norm2((const cplx *)v, n); }

/* ---------------------------------------------------------------- errors -- */
static __thread char g_err[512];
static void set_err(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); }
const char *qo_last_error(void) { return g_err; }

/* ------------------------------------------------------- enumeration ----- */
static int64_t binom64(int n, int k)
{
    if (k < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    int64_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

/* basis.jl:229-250: all L-bit words of popcount w in ascending order
 * (Gosper's hack), special-casing w = 0 and w = L.  w < 0 enumerates every
 * L-bit word (basis.jl:667-669).  Returns the count; out may be NULL. */
int64_t qo_fixed_weight_states(int L, int w, uint64_t *out)
{
    if (w < 0) {
        int64_t n = (int64_t)1 << L;
        if (out)
            for (int64_t i = 0; i < n; i++) out[i] = (uint64_t)i;
        return n;
    }
    int64_t count = binom64(L, w);
    if (!out) return count;
    if (w == 0) { out[0] = 0; return count; }
    if (w == L) { out[0] = (L == 64) ? ~0ULL : ((1ULL << L) - 1); return count; }
    uint64_t state = (1ULL << w) - 1;
    for (int64_t i = 0; i < count; i++) {
        out[i] = state;
        if (i == count - 1) break;
        uint64_t lowest = state & (~state + 1);
        uint64_t next = state + lowest;
        state = next | ((next ^ state) >> (__builtin_ctzll(lowest) + 2));
    }
    return count;
}

static int64_t lookup_state(const uint64_t *states, int64_t n, uint64_t s)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (states[mid] < s) lo = mid + 1; else hi = mid;
    }
    return (lo < n && states[lo] == s) ? lo : -1;
}

/* ----------------------------------------------------------- sparse ----- */
typedef struct {
    int64_t m, n, nnz;
    int64_t *colptr; /* n+1 */
    int64_t *row;
    cplx *val;
} csc_t;

static void csc_free(csc_t *a)
{
    free(a->colptr); free(a->row); free(a->val);
    memset(a, 0, sizeof *a);
}

typedef struct { int64_t row, col; cplx v; } trip_t;
static int trip_cmp(const void *pa, const void *pb)
{
    const trip_t *a = pa, *b = pb;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    return 0;
}

/* SparseArrays.sparse(I, J, V, m, n) for duplicate-free triplets. */
static int csc_from_triplets(csc_t *out, trip_t *t, int64_t nt, int64_t m, int64_t n)
{
    qsort(t, (size_t)nt, sizeof *t, trip_cmp);
    out->m = m; out->n = n; out->nnz = nt;
    out->colptr = calloc((size_t)n + 1, sizeof(int64_t));
    out->row = malloc(sizeof(int64_t) * (size_t)(nt ? nt : 1));
    out->val = malloc(sizeof(cplx) * (size_t)(nt ? nt : 1));
    if (!out->colptr || !out->row || !out->val) return -1;
    for (int64_t i = 0; i < nt; i++) {
        out->colptr[t[i].col + 1]++;
        out->row[i] = t[i].row;
        out->val[i] = t[i].v;
    }
    for (int64_t j = 0; j < n; j++) out->colptr[j + 1] += out->colptr[j];
    return 0;
}

/* copy(adjoint(A)) as CSC. */
static int csc_adjoint(csc_t *out, const csc_t *a)
{
    out->m = a->n; out->n = a->m; out->nnz = a->nnz;
    out->colptr = calloc((size_t)out->n + 1, sizeof(int64_t));
    out->row = malloc(sizeof(int64_t) * (size_t)(a->nnz ? a->nnz : 1));
    out->val = malloc(sizeof(cplx) * (size_t)(a->nnz ? a->nnz : 1));
    if (!out->colptr || !out->row || !out->val) return -1;
    for (int64_t p = 0; p < a->nnz; p++) out->colptr[a->row[p] + 1]++;
    for (int64_t j = 0; j < out->n; j++) out->colptr[j + 1] += out->colptr[j];
    int64_t *next = malloc(sizeof(int64_t) * (size_t)(out->n + 1));
    memcpy(next, out->colptr, sizeof(int64_t) * (size_t)(out->n + 1));
    for (int64_t c = 0; c < a->n; c++)
        for (int64_t p = a->colptr[c]; p < a->colptr[c + 1]; p++) {
            int64_t dst = next[a->row[p]]++;
            out->row[dst] = c;
            out->val[dst] = c_conj(a->val[p]);
        }
    free(next);
    return 0;
}

/* SparseArrays spmatmul (Gustavson, dense accumulator): products are formed
 * as nzA * nzB, the first product at a row is assigned and later ones added,
 * in the order of B's column entries and then A's column entries; each output
 * column is emitted in ascending row order. */
static int csc_matmul(csc_t *out, const csc_t *a, const csc_t *b)
{
    int64_t cap = a->nnz + b->nnz + 16;
    out->m = a->m; out->n = b->n; out->nnz = 0;
    out->colptr = calloc((size_t)b->n + 1, sizeof(int64_t));
    out->row = malloc(sizeof(int64_t) * (size_t)cap);
    out->val = malloc(sizeof(cplx) * (size_t)cap);
    cplx *x = malloc(sizeof(cplx) * (size_t)(a->m ? a->m : 1));
    char *xb = calloc((size_t)(a->m ? a->m : 1), 1);
    int64_t *touched = malloc(sizeof(int64_t) * (size_t)(a->m ? a->m : 1));
    if (!out->colptr || !out->row || !out->val || !x || !xb || !touched) return -1;
    int64_t ip = 0;
    for (int64_t i = 0; i < b->n; i++) {
        out->colptr[i] = ip;
        int64_t nt = 0;
        for (int64_t jp = b->colptr[i]; jp < b->colptr[i + 1]; jp++) {
            cplx nzb = b->val[jp];
            int64_t j = b->row[jp];
            for (int64_t kp = a->colptr[j]; kp < a->colptr[j + 1]; kp++) {
                cplx nzc = c_mul(a->val[kp], nzb);
                int64_t k = a->row[kp];
                if (xb[k]) {
                    x[k] = c_add(x[k], nzc);
                } else {
                    x[k] = nzc;
                    xb[k] = 1;
                    touched[nt++] = k;
                }
            }
        }
        /* ascending rows */
        for (int64_t u = 1; u < nt; u++) {
            int64_t key = touched[u], v = u - 1;
            while (v >= 0 && touched[v] > key) { touched[v + 1] = touched[v]; v--; }
            touched[v + 1] = key;
        }
        if (ip + nt > cap) {
            cap = (ip + nt) * 2;
            out->row = realloc(out->row, sizeof(int64_t) * (size_t)cap);
            out->val = realloc(out->val, sizeof(cplx) * (size_t)cap);
        }
        for (int64_t u = 0; u < nt; u++) {
            out->row[ip] = touched[u];
            out->val[ip] = x[touched[u]];
            xb[touched[u]] = 0;
            ip++;
        }
    }
    out->colptr[b->n] = ip;
    out->nnz = ip;
    free(x); free(xb); free(touched);
    return 0;
}

/* SparseArrays.droptol!(A, tol): drop entries with abs(v) <= tol. */
static void csc_droptol(csc_t *a, double tol)
{
    int64_t w = 0;
    for (int64_t c = 0; c < a->n; c++) {
        int64_t start = a->colptr[c];
        a->colptr[c] = w;
        for (int64_t p = start; p < a->colptr[c + 1]; p++) {
            if (c_abs(a->val[p]) <= tol) continue;
            a->row[w] = a->row[p];
            a->val[w] = a->val[p];
            w++;
        }
    }
    a->colptr[a->n] = w;
    a->nnz = w;
}

/* --------------------------------------------------------- transforms --- */
enum { TR_TRANSLATE = 0, TR_PARITY, TR_FLIP, TR_PARITY_FLIP, TR_FLIP_A, TR_FLIP_B, TR_GENERAL };

/* TR_GENERAL: a site map of SpinBasisGeneral (basis.jl:325-338): the bit of
 * site `source` (1-based), inverted when flips[source - 1], moves to site
 * perm[source - 1]. */
typedef struct { int kind; int L; int a; const int *perm; const unsigned char *flips; } transform_t;

static uint64_t full_mask(int L) { return L == 64 ? ~0ULL : ((1ULL << L) - 1); }

/* basis.jl:689-735: site permutations act as state -> OR of 1 << (perm[site]-1);
 * translation by a: site -> mod1(site + a, L); parity: site -> L - site + 1;
 * flips xor with the full / odd-site / even-site masks. */
static uint64_t apply_transform(const transform_t *t, uint64_t s)
{
    int L = t->L;
    uint64_t out = 0;
    switch (t->kind) {
    case TR_TRANSLATE:
        for (int site = 1; site <= L; site++) {
            if (!(s & (1ULL << (site - 1)))) continue;
            int dst = ((site + t->a - 1) % L) + 1;
            out |= 1ULL << (dst - 1);
        }
        return out;
    case TR_PARITY:
    case TR_PARITY_FLIP:
        for (int site = 1; site <= L; site++) {
            if (!(s & (1ULL << (site - 1)))) continue;
            out |= 1ULL << (L - site);
        }
        if (t->kind == TR_PARITY_FLIP) out ^= full_mask(L);
        return out;
    case TR_FLIP:
        return s ^ full_mask(L);
    case TR_FLIP_A: {
        uint64_t m = 0;
        for (int site = 1; site <= L; site += 2) m |= 1ULL << (site - 1);
        return s ^ m;
    }
    case TR_FLIP_B: {
        uint64_t m = 0;
        for (int site = 2; site <= L; site += 2) m |= 1ULL << (site - 1);
        return s ^ m;
    }
    case TR_GENERAL:
        for (int source = 1; source <= L; source++) {
            int occupied = (s >> (source - 1)) & 1;
            if (t->flips[source - 1]) occupied = !occupied;
            if (occupied) out |= 1ULL << (t->perm[source - 1] - 1);
        }
        return out;
    }
    return s;
}

/* symmetry_basis.jl:57-108, on the image table of the map: img[i] = index
 * of the image of parent state i (-1: outside the parent states). */
static int cyclic_projector_img(csc_t *out, const int64_t *img, int64_t n, cplx lambda)
{
    char *visited = calloc((size_t)n, 1);
    int64_t cap = n, nt = 0, ncol = 0;
    trip_t *t = malloc(sizeof(trip_t) * (size_t)cap);
    int64_t *orbit = malloc(sizeof(int64_t) * (size_t)(n + 1));
    cplx *coef = malloc(sizeof(cplx) * (size_t)(n + 1));
    if (!visited || !t || !orbit || !coef) { set_err("out of memory"); return -1; }
    int rc = 0;
    for (int64_t start = 0; start < n; start++) {
        if (visited[start]) continue;
        int len = 0;
        int64_t cur = start;
        for (;;) {
            if (len >= n) { set_err("symmetry map produced a non-cyclic orbit"); rc = -1; goto done; }
            orbit[len++] = cur;
            visited[cur] = 1;
            int64_t nxt = img[cur];
            if (nxt < 0) { set_err("symmetry transformation leaves the selected particle sector"); rc = -1; goto done; }
            if (nxt == start) break;
            for (int q = 0; q < len; q++)
                if (orbit[q] == nxt) { set_err("symmetry map produced a non-cyclic orbit"); rc = -1; goto done; }
            cur = nxt;
        }
        cplx one = c_make(1.0, 0.0);
        coef[0] = one;
        for (int q = 1; q < len; q++) coef[q] = c_div(c_mul(coef[q - 1], one), lambda);
        cplx closure = c_mul(coef[len - 1], one);
        if (!c_isapprox(closure, c_mul(lambda, coef[0]), 2e-10, 2e-10)) continue;
        double nrm = generic_norm2(coef, len);
        for (int q = 0; q < len; q++) {
            t[nt].row = orbit[q];
            t[nt].col = ncol;
            t[nt].v = c_make(coef[q].re / nrm, coef[q].im / nrm);
            nt++;
        }
        ncol++;
    }
    rc = csc_from_triplets(out, t, nt, n, ncol);
done:
    free(visited); free(t); free(orbit); free(coef);
    return rc;
}

/* symmetry_basis.jl:38-55, on the image table of the map */
static int signed_permutation_img(csc_t *out, const int64_t *img, int64_t n)
{
    out->m = n; out->n = n; out->nnz = n;
    out->colptr = malloc(sizeof(int64_t) * (size_t)(n + 1));
    out->row = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    out->val = malloc(sizeof(cplx) * (size_t)(n ? n : 1));
    if (!out->colptr || !out->row || !out->val) { set_err("out of memory"); return -1; }
    for (int64_t c = 0; c < n; c++) {
        if (img[c] < 0) { set_err("symmetry transformation leaves the selected particle sector"); return -1; }
        out->colptr[c] = c;
        out->row[c] = img[c];
        out->val[c] = c_make(1.0, 0.0);
    }
    out->colptr[n] = n;
    return 0;
}

static int64_t *image_table(const uint64_t *states, int64_t n, const transform_t *tr)
{
    int64_t *img = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    if (!img) { set_err("out of memory"); return NULL; }
    for (int64_t i = 0; i < n; i++) img[i] = lookup_state(states, n, apply_transform(tr, states[i]));
    return img;
}

static int cyclic_projector(csc_t *out, const uint64_t *states, int64_t n,
                            const transform_t *tr, cplx lambda)
{
    int64_t *img = image_table(states, n, tr);
    if (!img) return -1;
    int rc = cyclic_projector_img(out, img, n, lambda);
    free(img);
    return rc;
}

static int signed_permutation(csc_t *out, const uint64_t *states, int64_t n, const transform_t *tr)
{
    int64_t *img = image_table(states, n, tr);
    if (!img) return -1;
    int rc = signed_permutation_img(out, img, n);
    free(img);
    return rc;
}

/* symmetry_basis.jl:110-213 (the permutation-like branch; the dense nullspace
 * fallback is never reached by spin-1/2 chain symmetries). */
static int intersect_eigenspace(csc_t *proj, const csc_t *sym, cplx lambda)
{
    if (proj->n == 0) return 0;
    csc_t pa = {0}, q = {0}, r = {0};
    if (csc_adjoint(&pa, proj) || csc_matmul(&q, &pa, sym) || csc_matmul(&r, &q, proj)) {
        set_err("out of memory"); return -1;
    }
    csc_droptol(&r, 5e-12);
    int64_t nr = r.n;
    int64_t *mapped = malloc(sizeof(int64_t) * (size_t)(nr ? nr : 1));
    cplx *phase = malloc(sizeof(cplx) * (size_t)(nr ? nr : 1));
    char *seen = calloc((size_t)(nr ? nr : 1), 1);
    int perm_like = 1;
    for (int64_t c = 0; c < nr; c++) {
        if (r.colptr[c + 1] - r.colptr[c] != 1) { perm_like = 0; break; }
        mapped[c] = r.row[r.colptr[c]];
        phase[c] = r.val[r.colptr[c]];
        double a = c_abs(phase[c]);
        if (!(fabs(a - 1.0) <= fmax(2e-10, 2e-10 * fmax(a, 1.0)))) { perm_like = 0; break; }
    }
    if (perm_like) {
        for (int64_t c = 0; c < nr; c++) {
            if (seen[mapped[c]]) { perm_like = 0; break; }
            seen[mapped[c]] = 1;
        }
    }
    if (!perm_like) {
        set_err("non-monomial symmetry intersection (dense fallback not restated)");
        csc_free(&pa); csc_free(&q); csc_free(&r);
        free(mapped); free(phase); free(seen);
        return -1;
    }
    memset(seen, 0, (size_t)(nr ? nr : 1));
    trip_t *t = malloc(sizeof(trip_t) * (size_t)(nr ? nr : 1));
    int64_t nt = 0, ncol = 0;
    int64_t orbit[130];
    cplx ph[130], coef[130];
    for (int64_t start = 0; start < nr; start++) {
        if (seen[start]) continue;
        int len = 0;
        int64_t cur = start;
        for (;;) {
            orbit[len] = cur;
            seen[cur] = 1;
            ph[len] = phase[cur];
            len++;
            int64_t nxt = mapped[cur];
            if (nxt == start) break;
            if (seen[nxt] || len >= 128) { perm_like = 0; break; }
            cur = nxt;
        }
        if (!perm_like) break;
        coef[0] = c_make(1.0, 0.0);
        for (int q2 = 1; q2 < len; q2++) coef[q2] = c_div(c_mul(coef[q2 - 1], ph[q2 - 1]), lambda);
        cplx closure = c_mul(coef[len - 1], ph[len - 1]);
        if (!c_isapprox(closure, c_mul(lambda, coef[0]), 2e-10, 2e-10)) continue;
        double nrm = generic_norm2(coef, len);
        for (int q2 = 0; q2 < len; q2++) {
            t[nt].row = orbit[q2];
            t[nt].col = ncol;
            t[nt].v = c_make(coef[q2].re / nrm, coef[q2].im / nrm);
            nt++;
        }
        ncol++;
    }
    csc_free(&pa); csc_free(&q); csc_free(&r);
    free(mapped); free(phase); free(seen);
    if (!perm_like) { free(t); set_err("non-monomial symmetry intersection"); return -1; }
    csc_t red = {0}, res = {0};
    csc_from_triplets(&red, t, nt, nr, ncol);
    free(t);
    if (csc_matmul(&res, proj, &red)) { set_err("out of memory"); return -1; }
    csc_droptol(&res, 5e-14);
    csc_free(&red);
    csc_free(proj);
    *proj = res;
    return 0;
}

/* ------------------------------------------------------------ sector ---- */
typedef struct qo_sector {
    int L, nup, pauli, reduced;
    int64_t npar;
    uint64_t *parent;   /* ascending parent states */
    csc_t P;            /* npar x ncol projector (reduced only) */
    uint64_t *labels;   /* reference representative label per column */
    int64_t ncol;
} qo_sector;

void qo_sector_free(qo_sector *s)
{
    if (!s) return;
    free(s->parent);
    free(s->labels);
    csc_free(&s->P);
    free(s);
}

#define QO_UNSET (-1000000)

/* basis.jl:622-777.  Block arguments equal to QO_UNSET are "nothing". */
qo_sector *qo_sector_build(int L, int nup, int pauli, int a, int kblock, int pblock,
                           int zblock, int pzblock, int zAblock, int zBblock)
{
    if (L < 1 || L > 63) { set_err("L must be between 1 and 63"); return NULL; }
    if (nup != QO_UNSET && (nup < 0 || nup > L)) { set_err("nup must lie between 0 and L"); return NULL; }
    if (nup != QO_UNSET && L > 40 && binom64(L, nup) > (int64_t)1 << 31) {
        set_err("sector too large for the oracle"); return NULL;
    }
    if (nup == QO_UNSET && L > 30) { set_err("sector too large for the oracle"); return NULL; }
    qo_sector *s = calloc(1, sizeof *s);
    s->L = L; s->nup = nup; s->pauli = pauli;
    int w = nup == QO_UNSET ? -1 : nup;
    s->npar = qo_fixed_weight_states(L, w, NULL);
    s->parent = malloc(sizeof(uint64_t) * (size_t)s->npar);
    if (!s->parent) { set_err("out of memory"); free(s); return NULL; }
    qo_fixed_weight_states(L, w, s->parent);
    int requested = kblock != QO_UNSET || pblock != QO_UNSET || zblock != QO_UNSET ||
                    pzblock != QO_UNSET || zAblock != QO_UNSET || zBblock != QO_UNSET;
    if (!requested) {
        s->reduced = 0;
        s->ncol = s->npar;
        s->labels = malloc(sizeof(uint64_t) * (size_t)s->npar);
        memcpy(s->labels, s->parent, sizeof(uint64_t) * (size_t)s->npar);
        return s;
    }
    if (!(a > 0 && L % a == 0)) { set_err("a must be a positive divisor of L"); qo_sector_free(s); return NULL; }
    int order = L / a;
    int have = 0;
    if (kblock != QO_UNSET) {
        int m = ((kblock % order) + order) % order;
        /* Julia: cis(2pi * momentum / order), 2pi rounded to Float64 */
        double ang = (6.283185307179586 * (double)m) / (double)order;
        cplx lam = c_make(cos(ang), sin(ang));
        transform_t tr = {TR_TRANSLATE, L, a};
        if (cyclic_projector(&s->P, s->parent, s->npar, &tr, lam)) { qo_sector_free(s); return NULL; }
        have = 1;
    }
    if (pblock != QO_UNSET && kblock != QO_UNSET) {
        int m = ((kblock % order) + order) % order;
        if (!(m == 0 || (order % 2 == 0 && m == order / 2))) {
            set_err("pblock can be combined with kblock only at k=0 or pi");
            qo_sector_free(s); return NULL;
        }
    }
    int values[5] = {pblock, zblock, pzblock, zAblock, zBblock};
    int kinds[5] = {TR_PARITY, TR_FLIP, TR_PARITY_FLIP, TR_FLIP_A, TR_FLIP_B};
    for (int i = 0; i < 5; i++) {
        if (values[i] == QO_UNSET) continue;
        if (values[i] != 1 && values[i] != -1) { set_err("block must be +1 or -1"); qo_sector_free(s); return NULL; }
        transform_t tr = {kinds[i], L, a};
        cplx lam = c_make((double)values[i], 0.0);
        if (!have) {
            if (cyclic_projector(&s->P, s->parent, s->npar, &tr, lam)) { qo_sector_free(s); return NULL; }
            have = 1;
        } else {
            csc_t S = {0};
            if (signed_permutation(&S, s->parent, s->npar, &tr)) { csc_free(&S); qo_sector_free(s); return NULL; }
            int rc = intersect_eigenspace(&s->P, &S, lam);
            csc_free(&S);
            if (rc) { qo_sector_free(s); return NULL; }
        }
    }
    if (s->P.n == 0) { set_err("the requested symmetry sector is empty"); qo_sector_free(s); return NULL; }
    s->reduced = 1;
    s->ncol = s->P.n;
    /* symmetry_basis.jl:215-232 */
    s->labels = malloc(sizeof(uint64_t) * (size_t)s->ncol);
    for (int64_t c = 0; c < s->ncol; c++) {
        int64_t p0 = s->P.colptr[c], best = p0;
        for (int64_t p = p0 + 1; p < s->P.colptr[c + 1]; p++)
            if (c_abs(s->P.val[p]) > c_abs(s->P.val[best])) best = p;
        s->labels[c] = s->parent[s->P.row[best]];
    }
    return s;
}

/* basis.jl:282-304 (_normalize_general_spin_map): one entry per site; maps
 * holding a zero or a negative entry are QuSpin's zero-based labels (a
 * negative entry -v flips the site and sends it to site v), otherwise
 * Julia's one-based labels (negative = flipped).  0 or an error message. */
static int normalize_general_map(const int *values, int L, int *perm, unsigned char *flips)
{
    int zero_based = 0;
    for (int i = 0; i < L; i++)
        if (values[i] <= 0) zero_based = 1;
    for (int i = 0; i < L; i++) {
        int v = values[i];
        if (zero_based) {
            perm[i] = v >= 0 ? v + 1 : -v;
            flips[i] = v < 0;
        } else {
            perm[i] = v < 0 ? -v : v;
            flips[i] = v < 0;
        }
    }
    char *hit = calloc((size_t)L + 1, 1);
    int ok = 1;
    for (int i = 0; i < L; i++) {
        if (perm[i] < 1 || perm[i] > L || hit[perm[i]]) ok = 0;
        else hit[perm[i]] = 1;
    }
    free(hit);
    if (!ok) { set_err("a general symmetry map must be a site permutation"); return -1; }
    return 0;
}

static int64_t gcd64(int64_t a, int64_t b) { while (b) { int64_t t = a % b; a = b; b = t; } return a; }

/* basis.jl:306-323: lcm over cycles of the cycle length, doubled when the
 * cycle flips an odd number of sites */
static int general_map_order(const int *perm, const unsigned char *flips, int L)
{
    char *visited = calloc((size_t)L, 1);
    int64_t order = 1;
    for (int start = 0; start < L; start++) {
        if (visited[start]) continue;
        int cur = start, len = 0, parity = 0;
        while (!visited[cur]) {
            visited[cur] = 1;
            len++;
            parity ^= flips[cur];
            cur = perm[cur] - 1;
        }
        int64_t c = parity ? 2 * len : len;
        order = order / gcd64(order, c) * c;
    }
    free(visited);
    return (int)order;
}

/* basis.jl:530-620 (SpinBasisGeneral, L <= 63, spin 1/2): the parent is the
 * nup sector (or every state), and each block in the given order adds one
 * stage: the cyclic projector for the first, P'SP intersections after it,
 * with eigenvalue cis(2 pi mod(q, order) / order).  maps: nblocks x L. */
qo_sector *qo_sector_build_general(int L, int nup, int pauli, int nblocks, const int *maps, const int *qnumbers)
{
    if (L < 1 || L > 30) { set_err("sector too large for the oracle"); return NULL; }
    if (nup != QO_UNSET && (nup < 0 || nup > L)) { set_err("nup must lie between 0 and L"); return NULL; }
    qo_sector *s = calloc(1, sizeof *s);
    s->L = L; s->nup = nup; s->pauli = pauli;
    int w = nup == QO_UNSET ? -1 : nup;
    s->npar = qo_fixed_weight_states(L, w, NULL);
    s->parent = malloc(sizeof(uint64_t) * (size_t)s->npar);
    qo_fixed_weight_states(L, w, s->parent);
    if (nblocks == 0) {
        s->reduced = 0;
        s->ncol = s->npar;
        s->labels = malloc(sizeof(uint64_t) * (size_t)s->npar);
        memcpy(s->labels, s->parent, sizeof(uint64_t) * (size_t)s->npar);
        return s;
    }
    int *perm = malloc(sizeof(int) * (size_t)L);
    unsigned char *flips = malloc((size_t)L);
    for (int bi = 0; bi < nblocks; bi++) {
        if (normalize_general_map(maps + (size_t)bi * L, L, perm, flips)) goto fail;
        int order = general_map_order(perm, flips, L);
        int m = ((qnumbers[bi] % order) + order) % order;
        double ang = (6.283185307179586 * (double)m) / (double)order;
        cplx lam = c_make(cos(ang), sin(ang));
        transform_t tr = {TR_GENERAL, L, 1, perm, flips};
        if (bi == 0) {
            if (cyclic_projector(&s->P, s->parent, s->npar, &tr, lam)) goto fail;
        } else {
            csc_t S = {0};
            if (signed_permutation(&S, s->parent, s->npar, &tr)) { csc_free(&S); goto fail; }
            int rc = intersect_eigenspace(&s->P, &S, lam);
            csc_free(&S);
            if (rc) goto fail;
        }
    }
    free(perm); free(flips);
    if (s->P.n == 0) { set_err("the requested symmetry sector is empty"); qo_sector_free(s); return NULL; }
    s->reduced = 1;
    s->ncol = s->P.n;
    /* symmetry_basis.jl:215-232 */
    s->labels = malloc(sizeof(uint64_t) * (size_t)s->ncol);
    for (int64_t c = 0; c < s->ncol; c++) {
        int64_t p0 = s->P.colptr[c], best = p0;
        for (int64_t p = p0 + 1; p < s->P.colptr[c + 1]; p++)
            if (c_abs(s->P.val[p]) > c_abs(s->P.val[best])) best = p;
        s->labels[c] = s->parent[s->P.row[best]];
    }
    return s;
fail:
    free(perm); free(flips);
    qo_sector_free(s);
    return NULL;
}

int64_t qo_sector_ncols(const qo_sector *s) { return s->ncol; }
int64_t qo_sector_nparent(const qo_sector *s) { return s->npar; }
int64_t qo_sector_nnz(const qo_sector *s) { return s->reduced ? s->P.nnz : 0; }
int qo_sector_reduced(const qo_sector *s) { return s->reduced; }
void qo_sector_parent_states(const qo_sector *s, uint64_t *out)
{
    memcpy(out, s->parent, sizeof(uint64_t) * (size_t)s->npar);
}
void qo_sector_labels(const qo_sector *s, uint64_t *out)
{
    memcpy(out, s->labels, sizeof(uint64_t) * (size_t)s->ncol);
}
/* CSC export, 0-based; values as interleaved (re, im). */
void qo_sector_projector(const qo_sector *s, int64_t *colptr, int64_t *row, double *val)
{
    if (!s->reduced) return;
    memcpy(colptr, s->P.colptr, sizeof(int64_t) * (size_t)(s->P.n + 1));
    memcpy(row, s->P.row, sizeof(int64_t) * (size_t)s->P.nnz);
    memcpy(val, s->P.val, sizeof(cplx) * (size_t)s->P.nnz);
}

/* discrete_basis.jl:2386-2413: normalization(basis, label) = inv(abs2(amp)),
 * amp = the P entry in the row of the label (argmax |.| of that row; rows of
 * a spin-1/2 projector hold one entry).  Also the integer orbit size D. */
void qo_sector_norms(const qo_sector *s, double *ref_norm, uint32_t *orbit)
{
    for (int64_t c = 0; c < s->ncol; c++) {
        if (!s->reduced) {
            if (ref_norm) ref_norm[c] = 1.0;
            if (orbit) orbit[c] = 1;
            continue;
        }
        int64_t row = lookup_state(s->parent, s->npar, s->labels[c]);
        cplx amp = c_make(0, 0);
        for (int64_t p = s->P.colptr[c]; p < s->P.colptr[c + 1]; p++)
            if (s->P.row[p] == row) amp = s->P.val[p];
        if (ref_norm) ref_norm[c] = 1.0 / c_abs2(amp);
        if (orbit) orbit[c] = (uint32_t)(s->P.colptr[c + 1] - s->P.colptr[c]);
    }
}

/* ------------------------------------------------------- wide sector --- */
/* basis.jl:361-445 (_wide_spin_general_basis) and basis.jl:530-620: the same
 * stages for chains of up to 256 sites (the reference's UInt256 states),
 * states as four 64-bit words (site i, 1-based, is bit (i - 1) % 64 of word
 * (i - 1) / 64), the parent the nup sector in ascending order. */
typedef struct { uint64_t w[4]; } wstate;

static int ws_less(const wstate *a, const wstate *b)
{
    for (int k = 3; k >= 0; k--)
        if (a->w[k] != b->w[k]) return a->w[k] < b->w[k];
    return 0;
}

static int ws_eq(const wstate *a, const wstate *b)
{
    return a->w[0] == b->w[0] && a->w[1] == b->w[1] && a->w[2] == b->w[2] && a->w[3] == b->w[3];
}

static int64_t ws_lookup(const wstate *states, int64_t n, const wstate *s)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (ws_less(&states[mid], s)) lo = mid + 1; else hi = mid;
    }
    return (lo < n && ws_eq(&states[lo], s)) ? lo : -1;
}

/* basis.jl:325-338: the (inverted) bit of each source site moves to its target */
static wstate ws_apply_general(int L, const int *perm, const unsigned char *flips, const wstate *s)
{
    wstate out = {{0, 0, 0, 0}};
    for (int source = 1; source <= L; source++) {
        int occupied = (int)((s->w[(source - 1) >> 6] >> ((source - 1) & 63)) & 1);
        if (flips[source - 1]) occupied = !occupied;
        if (occupied) {
            int t = perm[source - 1] - 1;
            out.w[t >> 6] |= 1ULL << (t & 63);
        }
    }
    return out;
}

typedef struct qo_wsector {
    int L, nup, pauli, reduced;
    int64_t npar, ncol;
    wstate *parent;     /* ascending parent states */
    csc_t P;
    wstate *labels;
} qo_wsector;

void qo_wsector_free(qo_wsector *s)
{
    if (!s) return;
    free(s->parent);
    free(s->labels);
    csc_free(&s->P);
    free(s);
}

/* the particle sector of L sites and w particles in ascending order: the
 * occupied sites as an increasing tuple, advanced in colexicographic order */
static int64_t ws_fixed_weight(int L, int w, wstate *out, int64_t cap)
{
    int pos[256];
    for (int i = 0; i < w; i++) pos[i] = i;
    int64_t n = 0;
    for (;;) {
        if (n >= cap) return -1;
        wstate st = {{0, 0, 0, 0}};
        for (int i = 0; i < w; i++) st.w[pos[i] >> 6] |= 1ULL << (pos[i] & 63);
        out[n++] = st;
        int i = 0;
        while (i < w && pos[i] + 1 == (i + 1 < w ? pos[i + 1] : L)) i++;
        if (i == w) break;
        pos[i]++;
        for (int j = 0; j < i; j++) pos[j] = j;
    }
    return n;
}

qo_wsector *qo_wsector_build(int L, int nup, int pauli, int nblocks, const int *maps, const int *qnumbers)
{
    if (L < 1 || L > 256) { set_err("L must be between 1 and 256"); return NULL; }
    if (nup < 0 || nup > L) { set_err("nup must lie between 0 and L"); return NULL; }
    double cnt = 1;
    for (int i = 0; i < nup; i++) cnt = cnt * (double)(L - i) / (double)(i + 1);
    if (cnt > (double)(1 << 24)) { set_err("sector too large for the oracle"); return NULL; }
    qo_wsector *s = calloc(1, sizeof *s);
    s->L = L; s->nup = nup; s->pauli = pauli;
    int64_t cap = (int64_t)(cnt + 0.5) + 1;
    s->parent = malloc(sizeof(wstate) * (size_t)cap);
    s->npar = ws_fixed_weight(L, nup, s->parent, cap);
    if (s->npar < 0) { set_err("particle sector count"); qo_wsector_free(s); return NULL; }
    const int64_t n = s->npar;
    if (nblocks == 0) {
        s->reduced = 0;
        s->ncol = n;
        s->labels = malloc(sizeof(wstate) * (size_t)n);
        memcpy(s->labels, s->parent, sizeof(wstate) * (size_t)n);
        return s;
    }
    int *perm = malloc(sizeof(int) * (size_t)L);
    unsigned char *flips = malloc((size_t)L);
    int64_t *img = malloc(sizeof(int64_t) * (size_t)n);
    for (int bi = 0; bi < nblocks; bi++) {
        if (normalize_general_map(maps + (size_t)bi * L, L, perm, flips)) goto fail;
        int order = general_map_order(perm, flips, L);
        int m = ((qnumbers[bi] % order) + order) % order;
        double ang = (6.283185307179586 * (double)m) / (double)order;
        cplx lam = c_make(cos(ang), sin(ang));
        for (int64_t i = 0; i < n; i++) {
            wstate im = ws_apply_general(L, perm, flips, &s->parent[i]);
            img[i] = ws_lookup(s->parent, n, &im);
        }
        if (bi == 0) {
            if (cyclic_projector_img(&s->P, img, n, lam)) goto fail;
        } else {
            csc_t S = {0};
            if (signed_permutation_img(&S, img, n)) { csc_free(&S); goto fail; }
            int rc = intersect_eigenspace(&s->P, &S, lam);
            csc_free(&S);
            if (rc) goto fail;
        }
    }
    free(perm); free(flips); free(img);
    if (s->P.n == 0) { set_err("the requested symmetry sector is empty"); qo_wsector_free(s); return NULL; }
    s->reduced = 1;
    s->ncol = s->P.n;
    /* symmetry_basis.jl:215-232 */
    s->labels = malloc(sizeof(wstate) * (size_t)s->ncol);
    for (int64_t c = 0; c < s->ncol; c++) {
        int64_t p0 = s->P.colptr[c], best = p0;
        for (int64_t p = p0 + 1; p < s->P.colptr[c + 1]; p++)
            if (c_abs(s->P.val[p]) > c_abs(s->P.val[best])) best = p;
        s->labels[c] = s->parent[s->P.row[best]];
    }
    return s;
fail:
    free(perm); free(flips); free(img);
    qo_wsector_free(s);
    return NULL;
}

int64_t qo_wsector_ncols(const qo_wsector *s) { return s->ncol; }
int64_t qo_wsector_nparent(const qo_wsector *s) { return s->npar; }
int64_t qo_wsector_nnz(const qo_wsector *s) { return s->reduced ? s->P.nnz : 0; }
int qo_wsector_reduced(const qo_wsector *s) { return s->reduced; }
void qo_wsector_parent_states(const qo_wsector *s, uint64_t *out)
{
    memcpy(out, s->parent, sizeof(wstate) * (size_t)s->npar);
}
void qo_wsector_labels(const qo_wsector *s, uint64_t *out)
{
    memcpy(out, s->labels, sizeof(wstate) * (size_t)s->ncol);
}
void qo_wsector_projector(const qo_wsector *s, int64_t *colptr, int64_t *row, double *val)
{
    if (!s->reduced) return;
    memcpy(colptr, s->P.colptr, sizeof(int64_t) * (size_t)(s->P.n + 1));
    memcpy(row, s->P.row, sizeof(int64_t) * (size_t)s->P.nnz);
    memcpy(val, s->P.val, sizeof(cplx) * (size_t)s->P.nnz);
}
/* discrete_basis.jl:2386-2413, as qo_sector_norms */
void qo_wsector_norms(const qo_wsector *s, double *ref_norm, uint32_t *orbit)
{
    for (int64_t c = 0; c < s->ncol; c++) {
        if (!s->reduced) {
            if (ref_norm) ref_norm[c] = 1.0;
            if (orbit) orbit[c] = 1;
            continue;
        }
        int64_t row = ws_lookup(s->parent, s->npar, &s->labels[c]);
        cplx amp = c_make(0, 0);
        for (int64_t p = s->P.colptr[c]; p < s->P.colptr[c + 1]; p++)
            if (s->P.row[p] == row) amp = s->P.val[p];
        if (ref_norm) ref_norm[c] = 1.0 / c_abs2(amp);
        if (orbit) orbit[c] = (uint32_t)(s->P.colptr[c + 1] - s->P.colptr[c]);
    }
}

/* ---------------------------------------------------------- operator ---- */
typedef struct {
    char op[16];
    int nops;
    int sites[16];   /* 1-based */
    cplx coef;
} qo_coupling;

typedef struct qo_op {
    const qo_sector *sec;
    int pauli;
    int ncoup;
    qo_coupling *c;       /* action couplings, term-major in first-seen order */
    cplx *extracted;      /* ncol, extracted diagonal (check_symm path) */
    cplx *diagonal;       /* ncol, user diagonal */
} qo_op;

/* operators.jl:343-379 on one local operator. */
static int apply_local(int pauli, uint64_t *state, char op, int site, cplx *factor)
{
    uint64_t mask = 1ULL << (site - 1);
    int occ = (*state & mask) != 0;
    switch (op) {
    case 'I': *factor = c_make(1, 0); return 1;
    case 'z': { double sc = pauli ? 1.0 : 0.5; *factor = c_make(occ ? sc : -sc, 0); return 1; }
    case '+':
        if (occ) return 0;
        *factor = c_make(pauli ? 2.0 : 1.0, 0); *state |= mask; return 1;
    case '-':
        if (!occ) return 0;
        *factor = c_make(pauli ? 2.0 : 1.0, 0); *state &= ~mask; return 1;
    case 'x': { double sc = pauli ? 1.0 : 0.5; *factor = c_make(sc, 0); *state ^= mask; return 1; }
    case 'y': { double sc = pauli ? 1.0 : 0.5; *factor = c_make(0, occ ? -sc : sc); *state ^= mask; return 1; }
    }
    return -1;
}

/* Apply one coupling to a state, operators right to left (operators.jl:2569-2580). */
static int apply_coupling(int pauli, const qo_coupling *c, uint64_t s, uint64_t *out, cplx *amp)
{
    cplx a = c->coef;
    for (int i = c->nops - 1; i >= 0; i--) {
        cplx f;
        int r = apply_local(pauli, &s, c->op[i], c->sites[i], &f);
        if (r <= 0) return r;
        a = c_mul(a, f);
    }
    *out = s;
    *amp = a;
    return 1;
}

static int is_structurally_diagonal(const char *op)
{
    for (; *op; op++)
        if (*op != 'I' && *op != 'z' && *op != 'n' && *op != '|') return 0;
    return 1;
}

void qo_op_free(qo_op *o)
{
    if (!o) return;
    free(o->c); free(o->extracted); free(o->diagonal); free(o);
}

/* operators.jl:2492-2552 + 214-263.  Terms: n_terms operator strings; term t
 * has ncoup[t] couplings; sites is the flat list of 1-based sites, len(op)
 * per coupling.  check_symm selects the diagonal-extraction path
 * (operators.jl:2500-2507). diag may be NULL. */
qo_op *qo_op_create(const qo_sector *sec, int n_terms, const char **ops, const int *ncoup,
                    const double *coef, const int *sites, int check_symm, const double *diag)
{
    qo_op *o = calloc(1, sizeof *o);
    o->sec = sec;
    o->pauli = sec->pauli;
    int total = 0;
    for (int t = 0; t < n_terms; t++) total += ncoup[t];
    qo_coupling *raw = calloc((size_t)(total ? total : 1), sizeof *raw);
    int k = 0, sp = 0;
    for (int t = 0; t < n_terms; t++) {
        int len = (int)strlen(ops[t]);
        if (len < 1 || len > 15) { set_err("operator string length unsupported"); free(raw); free(o); return NULL; }
        for (int j = 0; j < ncoup[t]; j++) {
            qo_coupling *c = &raw[k++];
            strcpy(c->op, ops[t]);
            c->nops = len;
            for (int i = 0; i < len; i++) {
                c->sites[i] = sites[sp++];
                if (c->sites[i] < 1 || c->sites[i] > sec->L) { set_err("site outside 1:L"); free(raw); free(o); return NULL; }
                char ch = c->op[i];
                if (ch != 'I' && ch != 'z' && ch != '+' && ch != '-' && ch != 'x' && ch != 'y') {
                    set_err("unsupported spin operator"); free(raw); free(o); return NULL;
                }
            }
            c->coef = c_make(coef[2 * (k - 1)], coef[2 * (k - 1) + 1]);
        }
    }
    /* _consolidate_terms: key (op, sites) in first-seen order, summed, zeros
     * dropped, then grouped by operator string in first-seen order. */
    qo_coupling *keyed = calloc((size_t)(total ? total : 1), sizeof *keyed);
    int nk = 0;
    for (int i = 0; i < total; i++) {
        int found = -1;
        for (int j = 0; j < nk; j++)
            if (!strcmp(keyed[j].op, raw[i].op) && !memcmp(keyed[j].sites, raw[i].sites, sizeof(int) * 16)) { found = j; break; }
        if (found >= 0) keyed[found].coef = c_add(keyed[found].coef, raw[i].coef);
        else keyed[nk++] = raw[i];
    }
    free(raw);
    o->c = calloc((size_t)(nk ? nk : 1), sizeof *o->c);
    char done[4096] = {0};
    int nc = 0;
    for (int i = 0; i < nk && i < 4096; i++) {
        if (done[i]) continue;
        for (int j = i; j < nk; j++) {
            if (done[j] || strcmp(keyed[j].op, keyed[i].op)) continue;
            done[j] = 1;
            if (keyed[j].coef.re == 0 && keyed[j].coef.im == 0) continue;
            o->c[nc++] = keyed[j];
        }
    }
    free(keyed);
    o->ncoup = nc;
    o->extracted = calloc((size_t)sec->ncol, sizeof(cplx));
    o->diagonal = calloc((size_t)sec->ncol, sizeof(cplx));
    if (diag)
        for (int64_t i = 0; i < sec->ncol; i++) o->diagonal[i] = c_make(diag[2 * i], diag[2 * i + 1]);
    int extract = check_symm || !sec->reduced;
    if (extract) {
        /* operators.jl:2372-2429: parent diagonal of the diagonal terms, then
         * sum |P_ij|^2 d_i per column. */
        cplx *pd = calloc((size_t)sec->npar, sizeof(cplx));
        int keep = 0;
        for (int i = 0; i < nc; i++) {
            if (!is_structurally_diagonal(o->c[i].op)) { o->c[keep++] = o->c[i]; continue; }
            for (int64_t col = 0; col < sec->npar; col++) {
                uint64_t out;
                cplx amp;
                if (apply_coupling(o->pauli, &o->c[i], sec->parent[col], &out, &amp) == 1)
                    pd[col] = c_add(pd[col], amp);
            }
        }
        o->ncoup = keep;
        if (sec->reduced) {
            for (int64_t c = 0; c < sec->ncol; c++)
                for (int64_t p = sec->P.colptr[c]; p < sec->P.colptr[c + 1]; p++) {
                    double w2 = c_abs2(sec->P.val[p]);
                    cplx d = pd[sec->P.row[p]];
                    o->extracted[c] = c_add(o->extracted[c], c_make(w2 * d.re, w2 * d.im));
                }
        } else {
            memcpy(o->extracted, pd, sizeof(cplx) * (size_t)sec->npar);
        }
        free(pd);
    }
    return o;
}

/* operators.jl:2554-2589 on a basis without symmetry: push form with lookup. */
static void apply_terms_plain(const qo_op *o, const uint64_t *states, int64_t n,
                              const cplx *x, cplx *y, cplx alpha, int transposed, int conjugated,
                              int nthreads)
{
    (void)nthreads;
    for (int i = 0; i < o->ncoup; i++) {
        const qo_coupling *c = &o->c[i];
        qo_coupling ca = *c;
        ca.coef = c_mul(alpha, c->coef);
        for (int64_t col = 0; col < n; col++) {
            uint64_t out;
            cplx amp;
            if (apply_coupling(o->pauli, &ca, states[col], &out, &amp) != 1) continue;
            int64_t row = lookup_state(states, n, out);
            if (row < 0) continue;
            cplx tr = conjugated ? c_conj(amp) : amp;
            int64_t src = transposed ? row : col, dst = transposed ? col : row;
            cplx in = x[src];
            if (in.re == 0 && in.im == 0) continue;
            y[dst] = c_add(y[dst], c_mul(tr, in));
        }
    }
}

/* operators.jl:3147-3179 with operators.jl:2591-2636:
 * y = beta*y + alpha*(P'·H_parent·P x + (diag + extracted) .* x) */
int qo_op_matvec(const qo_op *o, const double *xin, double *yout, double a_re, double a_im,
                 double b_re, double b_im, int transposed, int conjugated)
{
    const qo_sector *s = o->sec;
    int64_t n = s->ncol;
    const cplx *x = (const cplx *)xin;
    cplx *y = (cplx *)yout;
    cplx alpha = c_make(a_re, a_im), beta = c_make(b_re, b_im);
    cplx *xc = malloc(sizeof(cplx) * (size_t)n);
    memcpy(xc, x, sizeof(cplx) * (size_t)n);
    for (int64_t i = 0; i < n; i++)
        y[i] = (beta.re == 0 && beta.im == 0) ? c_make(0, 0) : c_mul(beta, y[i]);
    if (s->reduced) {
        int ordinary = transposed == conjugated;
        cplx *lift = calloc((size_t)s->npar, sizeof(cplx));
        cplx *acted = calloc((size_t)s->npar, sizeof(cplx));
        for (int64_t c = 0; c < n; c++)
            for (int64_t p = s->P.colptr[c]; p < s->P.colptr[c + 1]; p++) {
                cplx pv = ordinary ? s->P.val[p] : c_conj(s->P.val[p]);
                lift[s->P.row[p]] = c_add(lift[s->P.row[p]], c_mul(pv, xc[c]));
            }
        apply_terms_plain(o, s->parent, s->npar, lift, acted, c_make(1, 0), transposed, conjugated, 1);
        for (int64_t c = 0; c < n; c++) {
            cplx acc = c_make(0, 0);
            for (int64_t p = s->P.colptr[c]; p < s->P.colptr[c + 1]; p++) {
                cplx pv = ordinary ? c_conj(s->P.val[p]) : s->P.val[p];
                acc = c_add(acc, c_mul(pv, acted[s->P.row[p]]));
            }
            y[c] = c_add(y[c], c_mul(alpha, acc));
        }
        free(lift); free(acted);
    } else {
        apply_terms_plain(o, s->parent, s->npar, xc, y, alpha, transposed, conjugated, 1);
    }
    for (int64_t i = 0; i < n; i++) {
        cplx d = c_add(o->diagonal[i], o->extracted[i]);
        if (conjugated) d = c_conj(d);
        y[i] = c_add(y[i], c_mul(alpha, c_mul(d, xc[i])));
    }
    free(xc);
    return 0;
}

int qo_op_ncouplings(const qo_op *o) { return o->ncoup; }

/* ------------------------------------------------------------ CSR ------- */
typedef struct qo_csr {
    int64_t n, nnz;
    int64_t *rowptr, *col;
    double *val;   /* real FP64 (the L=16 XXZ benchmark case is real) */
} qo_csr;

typedef struct { int64_t row, col; double v; } rtrip_t;
static int rtrip_cmp(const void *pa, const void *pb)
{
    const rtrip_t *a = pa, *b = pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    return 0;
}

/* operators.jl:381-442,540-576 for a real operator on a basis without
 * symmetry: off-diagonal triplets per term and coupling, the diagonal of the
 * 'z'/'I' terms accumulated separately and dropped where exactly zero,
 * duplicates summed, stored row-major (sparse_formats.jl:29-50). */
qo_csr *qo_csr_build(const qo_sector *sec, int n_terms, const char **ops, const int *ncoup,
                     const double *coef, const int *sites)
{
    if (sec->reduced) { set_err("CSR oracle covers bases without symmetry"); return NULL; }
    int64_t n = sec->npar;
    int64_t cap = n * 8 + 16, nt = 0;
    rtrip_t *t = malloc(sizeof(rtrip_t) * (size_t)cap);
    double *diag = calloc((size_t)n, sizeof(double));
    int sp = 0, k = 0;
    for (int term = 0; term < n_terms; term++) {
        int len = (int)strlen(ops[term]);
        int diagonal_term = 1;
        for (int i = 0; i < len; i++)
            if (ops[term][i] != 'I' && ops[term][i] != 'z') diagonal_term = 0;
        for (int j = 0; j < ncoup[term]; j++, k++) {
            qo_coupling c;
            memset(&c, 0, sizeof c);
            strcpy(c.op, ops[term]);
            c.nops = len;
            for (int i = 0; i < len; i++) c.sites[i] = sites[sp++];
            c.coef = c_make(coef[2 * k], 0.0);
            for (int64_t col = 0; col < n; col++) {
                uint64_t out;
                cplx amp;
                if (diagonal_term) {
                    /* operators.jl:418-442: every factor multiplies, no liveness */
                    if (apply_coupling(sec->pauli, &c, sec->parent[col], &out, &amp) == 1)
                        diag[col] += amp.re;
                    continue;
                }
                if (apply_coupling(sec->pauli, &c, sec->parent[col], &out, &amp) != 1) continue;
                int64_t row = lookup_state(sec->parent, n, out);
                if (row < 0) continue;
                if (nt == cap) { cap *= 2; t = realloc(t, sizeof(rtrip_t) * (size_t)cap); }
                t[nt].row = row; t[nt].col = col; t[nt].v = amp.re; nt++;
            }
        }
    }
    for (int64_t i = 0; i < n; i++) {
        if (diag[i] == 0.0) continue;
        if (nt == cap) { cap *= 2; t = realloc(t, sizeof(rtrip_t) * (size_t)cap); }
        t[nt].row = i; t[nt].col = i; t[nt].v = diag[i]; nt++;
    }
    free(diag);
    qsort(t, (size_t)nt, sizeof *t, rtrip_cmp);
    qo_csr *m = calloc(1, sizeof *m);
    m->n = n;
    m->rowptr = calloc((size_t)n + 1, sizeof(int64_t));
    m->col = malloc(sizeof(int64_t) * (size_t)(nt ? nt : 1));
    m->val = malloc(sizeof(double) * (size_t)(nt ? nt : 1));
    int64_t w = 0;
    for (int64_t i = 0; i < nt; i++) {
        if (i > 0 && t[i].row == t[i - 1].row && t[i].col == t[i - 1].col) {
            m->val[w - 1] += t[i].v;   /* duplicate (row, col): summed */
            continue;
        }
        m->col[w] = t[i].col;
        m->val[w] = t[i].v;
        m->rowptr[t[i].row + 1]++;
        w++;
    }
    for (int64_t i = 0; i < n; i++) m->rowptr[i + 1] += m->rowptr[i];
    m->nnz = w;
    free(t);
    return m;
}

int64_t qo_csr_nnz(const qo_csr *m) { return m->nnz; }
void qo_csr_free(qo_csr *m)
{
    if (!m) return;
    free(m->rowptr); free(m->col); free(m->val); free(m);
}

/* sparse_formats.jl:110-146: y = beta*y + alpha * A x, row-wise totals.
 * nthreads > 1 splits the rows across OpenMP threads (same per-row order). */
void qo_csr_matvec(const qo_csr *m, const double *x, double *y, double alpha, double beta, int nthreads)
{
    int64_t n = m->n;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t row = 0; row < n; row++) {
        double total = 0.0;
        for (int64_t p = m->rowptr[row]; p < m->rowptr[row + 1]; p++) total += m->val[p] * x[m->col[p]];
        y[row] = (beta == 0.0 ? 0.0 : beta * y[row]) + alpha * total;
    }
    (void)nthreads;
}

void qo_csr_export(const qo_csr *m, int64_t *rowptr, int64_t *col, double *val)
{
    memcpy(rowptr, m->rowptr, sizeof(int64_t) * (size_t)(m->n + 1));
    memcpy(col, m->col, sizeof(int64_t) * (size_t)m->nnz);
    memcpy(val, m->val, sizeof(double) * (size_t)m->nnz);
}

/* ------------------------------------------------------ CPU baseline ---- */
/* Times the reference's parent-space apply loop (operators.jl:2554-2589:
 * for each coupling, for each parent state: apply the local operators right
 * to left, look the image up, accumulate amplitude * x) over the first
 * n_sample parent states of the nup sector, single-threaded.  The lift P*x
 * and projection P'* (operators.jl:2604-2623) are not included, so the
 * extrapolated full-sector time is a lower bound on the reference's cost.
 * Returns seconds for the sample (best of `repeats`). */
#include <time.h>
/* ------------------------------------------- full QLO matvec baseline -- */
/* Open-addressing hash state -> parent index (the reference's Dict lookup,
 * basis.jl:671): linear probing, load <= 1/2, Fibonacci hashing. */
typedef struct {
    uint64_t *key;
    int64_t *val;
    uint64_t mask;
} hash_t;

static int hash_build(hash_t *h, const uint64_t *states, int64_t n)
{
    uint64_t cap = 1;
    while (cap < (uint64_t)(2 * n + 2)) cap <<= 1;
    h->mask = cap - 1;
    h->key = malloc(sizeof(uint64_t) * cap);
    h->val = malloc(sizeof(int64_t) * cap);
    if (!h->key || !h->val) return -1;
    for (uint64_t i = 0; i < cap; i++) h->key[i] = ~0ULL;
    for (int64_t i = 0; i < n; i++) {
        uint64_t p = (states[i] * 0x9E3779B97F4A7C15ULL) & h->mask;
        while (h->key[p] != ~0ULL) p = (p + 1) & h->mask;
        h->key[p] = states[i];
        h->val[p] = i;
    }
    return 0;
}

static inline int64_t hash_get(const hash_t *h, uint64_t s)
{
    uint64_t p = (s * 0x9E3779B97F4A7C15ULL) & h->mask;
    for (;;) {
        uint64_t k = h->key[p];
        if (k == s) return h->val[p];
        if (k == ~0ULL) return -1;
        p = (p + 1) & h->mask;
    }
}

static uint64_t rev_bits_u64(uint64_t u, int L)
{
    u = ((u >> 1) & 0x5555555555555555ULL) | ((u & 0x5555555555555555ULL) << 1);
    u = ((u >> 2) & 0x3333333333333333ULL) | ((u & 0x3333333333333333ULL) << 2);
    u = ((u >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((u & 0x0F0F0F0F0F0F0F0FULL) << 4);
    u = ((u >> 8) & 0x00FF00FF00FF00FFULL) | ((u & 0x00FF00FF00FF00FFULL) << 8);
    u = ((u >> 16) & 0x0000FFFF0000FFFFULL) | ((u & 0x0000FFFF0000FFFFULL) << 16);
    u = (u >> 32) | (u << 32);
    return u >> (64 - L);
}

static double bench_parent_apply(int L, int nup, int pauli, int n_terms, const char **ops, const int *ncoup,
                                 const double *coef, const int *sites, int64_t n_sample, int repeats, int threads)
{
    int64_t npar = qo_fixed_weight_states(L, nup, NULL);
    uint64_t *states = malloc(sizeof(uint64_t) * (size_t)npar);
    if (!states) return -1;
    qo_fixed_weight_states(L, nup, states);
    if (n_sample > npar) n_sample = npar;
    cplx *x = malloc(sizeof(cplx) * (size_t)npar);
    cplx *y = calloc((size_t)npar, sizeof(cplx));
    if (!x || !y) { free(states); free(x); free(y); return -1; }
    for (int64_t i = 0; i < npar; i++) x[i] = c_make(1.0 / (double)(1 + (i % 97)), 0.0);
    int total = 0;
    for (int t = 0; t < n_terms; t++) total += ncoup[t];
    qo_coupling *cs = calloc((size_t)(total ? total : 1), sizeof *cs);
    int k = 0, sp = 0;
    for (int t = 0; t < n_terms; t++) {
        int len = (int)strlen(ops[t]);
        for (int j = 0; j < ncoup[t]; j++, k++) {
            strcpy(cs[k].op, ops[t]);
            cs[k].nops = len;
            for (int i = 0; i < len; i++) cs[k].sites[i] = sites[sp++];
            cs[k].coef = c_make(coef[2 * k], coef[2 * k + 1]);
        }
    }
    hash_t hh;
    if (hash_build(&hh, states, npar)) { free(states); free(x); free(y); free(cs); return -1; }
    double *ts = malloc(sizeof(double) * (size_t)(repeats > 0 ? repeats : 1));
    for (int r = 0; r < repeats; r++) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        if (threads <= 1) {
            for (int i = 0; i < total; i++) {
                for (int64_t col = 0; col < n_sample; col++) {
                    uint64_t out;
                    cplx amp;
                    if (apply_coupling(pauli, &cs[i], states[col], &out, &amp) != 1) continue;
                    int64_t row = hash_get(&hh, out);
                    if (row < 0) continue;
                    cplx in = x[col];
                    if (in.re == 0 && in.im == 0) continue;
                    y[row] = c_add(y[row], c_mul(amp, in));
                }
            }
        } else {
            /* the same push loop over columns split across threads; the
               scattered additions into y are atomic */
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int64_t col = 0; col < n_sample; col++) {
                cplx in = x[col];
                if (in.re == 0 && in.im == 0) continue;
                for (int i = 0; i < total; i++) {
                    uint64_t out;
                    cplx amp;
                    if (apply_coupling(pauli, &cs[i], states[col], &out, &amp) != 1) continue;
                    int64_t row = hash_get(&hh, out);
                    if (row < 0) continue;
                    cplx p = c_mul(amp, in);
#pragma omp atomic
                    y[row].re += p.re;
#pragma omp atomic
                    y[row].im += p.im;
                }
            }
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        ts[r] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    }
    /* median of the repeats */
    for (int a = 1; a < repeats; a++)
        for (int b = a; b > 0 && ts[b - 1] > ts[b]; b--) { double t = ts[b]; ts[b] = ts[b - 1]; ts[b - 1] = t; }
    double med = repeats > 0 ? (repeats % 2 ? ts[repeats / 2] : 0.5 * (ts[repeats / 2 - 1] + ts[repeats / 2])) : 0.0;
    volatile double sink = y[0].re;
    (void)sink;
    free(states); free(x); free(y); free(cs); free(ts); free(hh.key); free(hh.val);
    return med;
}

/* Timed full product y = P' H_par P x of the reference's QLO in a k / p / z
 * sector (operators.jl:2591-2623: lift, parent apply with a Dict lookup per
 * (state, coupling), project), ComplexF64 throughout.  The projector columns
 * are built from the group orbits directly (ascending orbit minima, entries
 * sum_g conj(chi(g)) / norm; equal to the reference's stage construction up
 * to rounding), not timed.  Writes the seconds of each of `repeats` products
 * to times[] (times[3 r + 0..2] = lift, apply, project); returns the sector
 * dimension, or -1. */
int64_t qo_bench_qlo_matvec(int L, int nup, int pauli, int kblock, int pblock, int zblock, int n_terms,
                            const char **ops, const int *ncoup, const double *coef, const int *sites, int threads,
                            int repeats, double *times)
{
    int64_t npar = qo_fixed_weight_states(L, nup, NULL);
    uint64_t *states = malloc(sizeof(uint64_t) * (size_t)npar);
    if (!states) return -1;
    qo_fixed_weight_states(L, nup, states);
    hash_t h;
    if (hash_build(&h, states, npar)) return -1;
    const uint64_t full = L == 64 ? ~0ULL : ((1ULL << L) - 1);
    /* group: T^t o X_q, X_q in {1, P, Z, PZ} as requested */
    const int nT = kblock == QO_UNSET ? 1 : L;
    int npt = 1;
    int prev[4] = {0}, pfl[4] = {0};
    double pch[4] = {1, 0, 0, 0};
    if (pblock != QO_UNSET) { prev[npt] = 1; pfl[npt] = 0; pch[npt] = pblock; npt++; }
    if (zblock != QO_UNSET) {
        int m = npt;
        for (int q = 0; q < m; q++) { prev[npt] = prev[q]; pfl[npt] = 1; pch[npt] = pch[q] * zblock; npt++; }
    }
    /* column pointers of P: orbit by orbit in ascending order of the minimum */
    unsigned char *seen = calloc((size_t)npar, 1);
    int64_t *colptr = malloc(sizeof(int64_t) * (size_t)(npar + 1));
    int64_t *prow = malloc(sizeof(int64_t) * (size_t)npar);
    cplx *pval = malloc(sizeof(cplx) * (size_t)npar);
    if (!seen || !colptr || !prow || !pval) return -1;
    int64_t ncol = 0, nnz = 0;
    int G = nT * npt;
    int64_t *orb = malloc(sizeof(int64_t) * (size_t)G);
    cplx *acc = malloc(sizeof(cplx) * (size_t)G);
    for (int64_t i = 0; i < npar; i++) {
        if (seen[i]) continue;
        int no = 0;
        for (int t = 0; t < nT; t++)
            for (int q = 0; q < npt; q++) {
                uint64_t u = states[i];
                if (prev[q]) u = rev_bits_u64(u, L);
                if (pfl[q]) u = ~u & full;
                if (t) u = ((u << t) | (u >> (L - t))) & full;
                int64_t r = hash_get(&h, u);
                /* character of T^t X_q: exp(-2 pi i k t / L) * pch */
                double ang = kblock == QO_UNSET ? 0.0 : -2.0 * M_PI * (double)kblock * (double)t / (double)L;
                cplx ch = c_make(pch[q] * cos(ang), -pch[q] * sin(ang));   /* conj(chi) */
                int j = 0;
                while (j < no && orb[j] != r) j++;
                if (j == no) { orb[no] = r; acc[no] = c_make(0, 0); no++; }
                acc[j] = c_add(acc[j], ch);
            }
        double n2 = 0;
        for (int j = 0; j < no; j++) { seen[orb[j]] = 1; n2 += c_abs2(acc[j]); }
        if (n2 < 1e-18) continue;
        double inv = 1.0 / sqrt(n2);
        colptr[ncol] = nnz;
        /* rows ascending, as a CSC column */
        for (int a = 1; a < no; a++)
            for (int b = a; b > 0 && orb[b - 1] > orb[b]; b--) {
                int64_t tr = orb[b]; orb[b] = orb[b - 1]; orb[b - 1] = tr;
                cplx tc = acc[b]; acc[b] = acc[b - 1]; acc[b - 1] = tc;
            }
        for (int j = 0; j < no; j++) {
            prow[nnz] = orb[j];
            pval[nnz] = c_make(acc[j].re * inv, acc[j].im * inv);
            nnz++;
        }
        ncol++;
    }
    colptr[ncol] = nnz;
    free(seen); free(orb); free(acc);
    /* couplings */
    int total = 0;
    for (int t = 0; t < n_terms; t++) total += ncoup[t];
    qo_coupling *cs = calloc((size_t)(total ? total : 1), sizeof *cs);
    int k = 0, sp = 0;
    for (int t = 0; t < n_terms; t++) {
        int len = (int)strlen(ops[t]);
        for (int j = 0; j < ncoup[t]; j++, k++) {
            strcpy(cs[k].op, ops[t]);
            cs[k].nops = len;
            for (int i = 0; i < len; i++) cs[k].sites[i] = sites[sp++];
            cs[k].coef = c_make(coef[2 * k], coef[2 * k + 1]);
        }
    }
    cplx *x = malloc(sizeof(cplx) * (size_t)(ncol ? ncol : 1));
    cplx *y = malloc(sizeof(cplx) * (size_t)(ncol ? ncol : 1));
    cplx *lifted = malloc(sizeof(cplx) * (size_t)npar);
    cplx *acted = malloc(sizeof(cplx) * (size_t)npar);
    if (!x || !y || !lifted || !acted) return -1;
    for (int64_t c = 0; c < ncol; c++) x[c] = c_make(1.0 / (double)(1 + (c % 89)), 0.0);
    if (threads < 1) threads = 1;
    /* first touch of the work vectors outside the timed region */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < npar; i++) { lifted[i] = c_make(0, 0); acted[i] = c_make(0, 0); }
    for (int r = 0; r < repeats; r++) {
        struct timespec t0, t1, t2, t3;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        /* lifted = P x: columns own disjoint rows */
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t i = 0; i < npar; i++) lifted[i] = c_make(0, 0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
        for (int64_t c = 0; c < ncol; c++)
            for (int64_t q = colptr[c]; q < colptr[c + 1]; q++) lifted[prow[q]] = c_mul(pval[q], x[c]);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        /* acted = H_par lifted: the push loop with one Dict lookup per hit */
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t i = 0; i < npar; i++) acted[i] = c_make(0, 0);
        if (threads == 1) {
            for (int i = 0; i < total; i++)
                for (int64_t col = 0; col < npar; col++) {
                    cplx in = lifted[col];
                    if (in.re == 0 && in.im == 0) continue;
                    uint64_t out;
                    cplx amp;
                    if (apply_coupling(pauli, &cs[i], states[col], &out, &amp) != 1) continue;
                    int64_t row = hash_get(&h, out);
                    if (row < 0) continue;
                    acted[row] = c_add(acted[row], c_mul(amp, in));
                }
        } else {
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int64_t col = 0; col < npar; col++) {
                cplx in = lifted[col];
                if (in.re == 0 && in.im == 0) continue;
                for (int i = 0; i < total; i++) {
                    uint64_t out;
                    cplx amp;
                    if (apply_coupling(pauli, &cs[i], states[col], &out, &amp) != 1) continue;
                    int64_t row = hash_get(&h, out);
                    if (row < 0) continue;
                    cplx pr = c_mul(amp, in);
#pragma omp atomic
                    acted[row].re += pr.re;
#pragma omp atomic
                    acted[row].im += pr.im;
                }
            }
        }
        clock_gettime(CLOCK_MONOTONIC, &t2);
        /* y = P' acted */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
        for (int64_t c = 0; c < ncol; c++) {
            cplx sum = c_make(0, 0);
            for (int64_t q = colptr[c]; q < colptr[c + 1]; q++) sum = c_add(sum, c_mul(c_conj(pval[q]), acted[prow[q]]));
            y[c] = sum;
        }
        clock_gettime(CLOCK_MONOTONIC, &t3);
        times[3 * r + 0] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        times[3 * r + 1] = (double)(t2.tv_sec - t1.tv_sec) + 1e-9 * (double)(t2.tv_nsec - t1.tv_nsec);
        times[3 * r + 2] = (double)(t3.tv_sec - t2.tv_sec) + 1e-9 * (double)(t3.tv_nsec - t2.tv_nsec);
    }
    volatile double sink = ncol ? y[0].re : 0.0;
    (void)sink;
    free(states); free(h.key); free(h.val); free(colptr); free(prow); free(pval); free(cs);
    free(x); free(y); free(lifted); free(acted);
    return ncol;
}

/* Seconds of the reference's parent-space apply loop (operators.jl:2554-2589)
   over the first n_sample parent states, median of `repeats`: one thread. */
double qo_bench_parent_apply(int L, int nup, int pauli, int n_terms, const char **ops, const int *ncoup,
                             const double *coef, const int *sites, int64_t n_sample, int repeats)
{
    return bench_parent_apply(L, nup, pauli, n_terms, ops, ncoup, coef, sites, n_sample, repeats, 1);
}

/* The same on `threads` OpenMP threads (columns split, atomic scatter). */
double qo_bench_parent_apply_mt(int L, int nup, int pauli, int n_terms, const char **ops, const int *ncoup,
                                const double *coef, const int *sites, int64_t n_sample, int repeats, int threads)
{
    return bench_parent_apply(L, nup, pauli, n_terms, ops, ncoup, coef, sites, n_sample, repeats, threads);
}
