// qs_internal.h -- shared types of libqsgpu (host runtime + gfx950 kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qsgpu.h"

namespace qs {

using cd = std::complex<double>;

// ------------------------------------------------------------ errors ------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string &msg);
void set_last_error(const std::string &msg);

#define QS_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            ::qs::fail(QS_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// --------------------------------------------------------- device memory --
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            fail(QS_EOOM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T *get() const { return p; }
};

// fn(i) for i in [0, n) on up to 16 host threads (independent iterations:
// the small projected problems of many Krylov factorizations)
void parallel_for(size_t n, const std::function<void(size_t)> &fn);

// Device scratch allocated and freed in stream order (hipMallocAsync /
// hipFreeAsync on the solver's stream).  Unlike hipFree, releasing it does
// not wait for the whole device, so solvers on different streams (the
// serving loop's workers) keep overlapping; the default pool keeps up to a
// few GB of freed memory for reuse (stream_pool_setup).
void stream_pool_setup();
void stream_pool_trim();   // after a large batch: keep at most 8 GB cached
template <typename T>
struct SBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;
    SBuf() = default;
    SBuf(size_t count, hipStream_t st) { alloc(count, st); }
    SBuf(const SBuf &) = delete;
    SBuf &operator=(const SBuf &) = delete;
    ~SBuf() { release(); }
    void alloc(size_t count, hipStream_t st) {
        release();
        s = st;
        if (count == 0) return;
        stream_pool_setup();
        if (hipMallocAsync((void **)&p, count * sizeof(T), s) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            fail(QS_EOOM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFreeAsync(p, s);
        p = nullptr;
        n = 0;
    }
    T *get() const { return p; }
};

// Job descriptors of batched launches too many for the kernel arguments:
// written to pinned host memory and copied to a device table on the
// launch's stream.  Space is handed out in order and reused once that stream
// has drained (a full table waits for it, then starts over).
class JobTable {
public:
    JobTable() = default;
    JobTable(const JobTable &) = delete;
    JobTable &operator=(const JobTable &) = delete;
    ~JobTable();
    // stream-ordered copy of `bytes` from src; returns the device address
    const void *stage(const void *src, size_t bytes, hipStream_t s);

private:
    char *host_ = nullptr, *dev_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    hipStream_t s_ = nullptr;   // the stream every staged copy ran on
};

// -------------------------------------------------------------- groups ----
// A symmetry group element acts on an L-bit state as
//   g(u) = rotl_L( rev_L^r(u) ^ f , t * a )
// The group is enumerated as {T^t o X_p : t in [0, nT), p in point ops}.
constexpr int kMaxPoint = 16;
constexpr int kMaxRot = 64;

// A state of up to 256 sites (the reference's UInt256 basis integers):
// site i (0-based) is bit i % 64 of word i / 64.
constexpr int kMaxWideL = 256;
struct WState {
    uint64_t w[4] = {0, 0, 0, 0};
    bool operator<(const WState &o) const {
        for (int k = 3; k >= 0; k--)
            if (w[k] != o.w[k]) return w[k] < o.w[k];
        return false;
    }
    bool operator==(const WState &o) const {
        return w[0] == o.w[0] && w[1] == o.w[1] && w[2] == o.w[2] && w[3] == o.w[3];
    }
    bool operator!=(const WState &o) const { return !(*this == o); }
    WState operator^(const WState &o) const {
        WState r;
        for (int k = 0; k < 4; k++) r.w[k] = w[k] ^ o.w[k];
        return r;
    }
    WState operator&(const WState &o) const {
        WState r;
        for (int k = 0; k < 4; k++) r.w[k] = w[k] & o.w[k];
        return r;
    }
    bool bit(int i) const { return (w[i >> 6] >> (i & 63)) & 1; }
    void set(int i) { w[i >> 6] |= 1ULL << (i & 63); }
    int popcount() const {
        return __builtin_popcountll(w[0]) + __builtin_popcountll(w[1]) + __builtin_popcountll(w[2]) +
               __builtin_popcountll(w[3]);
    }
    bool zero() const { return (w[0] | w[1] | w[2] | w[3]) == 0; }
};

struct GroupHost {
    int L = 0, a = 1, nT = 1;
    int npoint = 1;
    int prev[kMaxPoint] = {0};
    uint64_t pflip[kMaxPoint] = {0};
    cd pchar[kMaxPoint];     // character of X_p
    cd lam[kMaxRot];         // character of T^t
    // general site maps (SpinBasisGeneral): every element as an explicit
    // map u -> perm(u) ^ c (bit i moves to bit perm[i]) with its character,
    // identity first; the (t, p) fields above are then unused
    bool general = false;
    std::vector<std::vector<int>> eperm;
    std::vector<uint64_t> ec;
    std::vector<cd> echar;
    std::vector<WState> ecw;    // wide bases: the xor constants on up to 256 sites
    int order() const { return general ? (int)eperm.size() : nT * npoint; }
};
constexpr int kMaxGeneral = 256;   // elements of a general-map group

// A SpinBasisGeneral symmetry block (basis.jl:282-338): the bit of source
// site i, inverted when bit i of `flip` is set, moves to site perm[i]
// (0-based); eigenvalue cis(2 pi q / order).
struct GenMap {
    std::vector<int> perm;
    uint64_t flip = 0;
    WState flipw;          // the same flips on up to 256 sites
    int order = 1, q = 0;
};

// Device-side copy passed by value in kernel arguments.
struct GroupDev {
    int L, a, nT, npoint;
    // L <= 32: funnel-shift window offset that reads the rotation T^t of a
    // periodic 64-bit copy of the state (see canonical32)
    int win[kMaxRot];
    int prev[kMaxPoint];
    uint64_t pflip[kMaxPoint];
    double pch_re[kMaxPoint], pch_im[kMaxPoint];
    double lam_re[kMaxRot], lam_im[kMaxRot];
};

// ------------------------------------------------------------- terms ------
// A compiled operator coupling acting on a basis state s:
//   alive iff (s & cmask) == cval;  s -> s ^ flip;
//   amplitude = A * (-1)^popcount(s & smask)
struct CTerm {
    uint64_t flip, cmask, cval, smask;
    cd A;
};

// The same on states of up to 256 sites (wide bases)
struct WTerm {
    WState flip, cmask, cval, smask;
    cd A;
};
struct WTermDev {
    uint64_t cmask[4], cval[4], smask[4];
    double are, aim;
};
struct WGroupDev {     // off-diagonal wide terms sharing one flip mask
    uint64_t flip[4];
    int t0, t1;
};
struct TermDev {
    uint64_t cmask, cval, smask;
    double are, aim;
};
struct GroupOfTerms {   // off-diagonal terms sharing one flip mask
    uint64_t flip;
    int t0, t1;
    int lo, hi;          // the two flipped bits when the mask has exactly two, else -1
};
// A flip group whose terms test and sign only its own two bits: the summed
// amplitude is a function of the 2-bit pattern (bit lo) | (bit hi) << 1.
struct HopDev {
    int lo, hi;
    uint32_t lowmask;    // bits below lo
    uint32_t between;    // bits strictly between lo and hi
    int lo65, hi65;      // binomial-table rows of lo and hi
    int nzmask;          // patterns with a nonzero amplitude (inside the particle sector)
    int sym;             // patterns 1 and 2 share one amplitude
    int top;             // hi is bit L-1
    int pad;
    double re[4], im[4];
};
// Diagonal terms A * (-1)^(s_i ^ s_{(i+d) mod L}) for every i in `mask`
// (d = 0: A * (-1)^s_i), summed as A * (n - 2 popcount(mask & (s ^ rotr(s, d)))).
struct DiagClass {
    uint64_t mask;
    int d, n;
    double re, im;
};
// Fast forms of one compiled term list (split_terms output).
struct FastTerms {
    bool hop_ok = false;           // every flip group is a two-bit hop
    std::vector<HopDev> hops;      // one per flip group, same order
    std::vector<DiagClass> dcls;
    std::vector<CTerm> drest;      // diagonal terms outside the classes
};
FastTerms compile_fast(const std::vector<CTerm> &diag, const std::vector<CTerm> &off,
                       const std::vector<GroupOfTerms> &groups, int L, int nup);

// Basis modes.
enum BasisMode : int {
    MODE_FULL = 0,   // no symmetry: column index = combinadic rank (or the state itself without nup)
    MODE_ZONLY = 1,  // spin inversion only at nup = L/2 (or no nup): canonical = top bit clear
    MODE_STORED = 2  // general group: sorted canonical states stored on the device
};

// Tables of the split matvec kernel for an implicit basis with a particle
// number (built on first use): H low bits, W = L - H high bits.
struct SplitBasis {
    int H = 0, W = 0;
    int nblocks = 0, maxblock = 0;
    int cls_off[34] = {}, cls_n[34] = {};
    DevBuf<uint32_t> blist;    // high words with a nonempty block, ascending
    DevBuf<uint32_t> bstart;   // first row of each of those blocks
    DevBuf<uint32_t> boff;     // first row of the block of each high word (2^W; ~0 if none)
    DevBuf<uint16_t> rank;     // rank of every H-bit word among the words of its popcount
    DevBuf<uint16_t> words;    // H-bit words by (popcount, value)
    std::vector<uint16_t> rank_h, words_h;   // host copies (lean-kernel tables)
    // block orders of the two-pass tiled form, one per (kA, kB), built on
    // first use: pass A tiles share the high-word bits >= kA, pass B tiles
    // share the high-word bits < kB (with the inverted tile under spin
    // inversion); each tile is padded to the same length with 0xFFFFFFFF
    struct Order {
        int kA = 0, kB = 0;
        int CA = 0, CB = 0;                   // entries per tile
        int nA = 0, nB = 0;                   // list lengths
        DevBuf<uint32_t> blA, bsA, blB, bsB;  // high word, first row
    };
    mutable std::mutex order_mutex;
    mutable std::vector<std::unique_ptr<Order>> orders;
};
// Hop groups of one term list classified for the split kernel: inside the
// low word, inside the high word, across the split (in this order).
struct SplitHops {
    int nlow = 0, nhigh = 0, ncross = 0;
    DevBuf<HopDev> d;
    // the lean kernel's neighbour tables (lean = they fit its limits): per
    // class position w, the low hops' neighbour rows (u8, 255 = none, stride
    // sl) and the cross hops' rows in the neighbour block with the word's bit
    // lo in bit 15 (u16, stride sc)
    bool lean = false;
    int sl = 0, sc = 0;
    DevBuf<uint8_t> lnb;
    DevBuf<uint16_t> cnb;
};
// The same groups dealt to the two passes of the tiled form for one (kA,
// kB): pass A takes every group whose upper bit lies below H + kA, pass B the
// rest, which must leave the high-word bits [0, kB) alone.
struct SplitPassHops {
    int kA = 0, kB = 0;
    bool ok = false, ok_adj = false;
    SplitHops a, b, a_adj, b_adj;
};

// The assembled sector matrix of an operator (plain products), SELL-64:
// rows in slices of 64, slice q padded to its longest row, column-major.
// Stored weights of assembled rows: per entry (re, im; im null for real
// weights) or, when the operator has few distinct weights, a 16-bit index
// per entry into a table of (re, im) pairs (dict non-null).
struct SellW {
    const double *re, *im;
    const uint16_t *idx;
    const double *dict;
};

struct SellHost {
    int64_t nslices = 0;
    uint64_t stored = 0;          // entries including padding
    uint64_t entries = 0;         // entries without padding
    bool cplx = false;            // complex weights
    DevBuf<uint64_t> sptr;        // first entry of each slice
    DevBuf<uint32_t> width;       // entries per row of each slice
    DevBuf<uint32_t> cols;
    DevBuf<double> wre, wim;      // per-entry weights (freed when the table form is kept)
    DevBuf<uint16_t> widx;        // table form: per-entry index into dict
    DevBuf<double> dict;          // table form: (re, im) pairs
    uint32_t ndict = 0;           // distinct weights in the table form
    SellW weights() const {
        if (dict.get()) return SellW{nullptr, nullptr, widx.get(), dict.get()};
        return SellW{wre.get(), cplx ? wim.get() : nullptr, nullptr, nullptr};
    }
};

struct BasisHost {
    int L = 0, nup = -1, pauli = 1, a = 1;
    int blocks[6];               // k, p, z, pz, zA, zB (QS_UNSET if absent)
    int mode = MODE_FULL;
    int is_complex = 0;          // sector phases are complex
    int use64 = 0;               // states need 64-bit words
    int64_t Ns = 0;
    int zchar = 1;               // MODE_ZONLY: character of the inversion
    GroupHost group;
    // MODE_STORED device arrays
    DevBuf<uint32_t> reps32;
    DevBuf<uint64_t> reps64;
    DevBuf<double> inv_nrm;      // 1/sqrt(orbit size)
    DevBuf<uint32_t> orbit;      // orbit size per column
    // open-addressing hash table canonical state -> column (linear probing,
    // load <= 1/2): for L <= 32 one u64 per slot, (state << 32) | column;
    // for L > 32 a key array and a column array.  Empty slots are all ones.
    DevBuf<uint64_t> hkeys;
    DevBuf<uint32_t> hvals;
    uint64_t hmask = 0;          // table size - 1
    int hmax_probe = 1;          // longest probe sequence of any stored state
    // search acceleration: every 2^sample_shift-th representative
    int sample_shift = 0;
    int64_t nsample = 0;
    DevBuf<uint32_t> samp32;
    DevBuf<uint64_t> samp64;
    // binomial table C(n, k) for n, k < 64 (device), combinadic ranking
    DevBuf<uint64_t> binom;
    // implicit modes, L <= 32: the column states themselves (u32), so rows
    // need no unranking
    bool states_stored = false;
    // stored mode, L <= 32: parent lookup table indexed by the parent rank
    // (combinadic rank, or the state itself without nup).  Entry = column |
    // (group element t * npoint + p mapping the state to its orbit minimum)
    // << ptab_colbits; all ones = the state's orbit is not in the sector.
    DevBuf<uint32_t> ptab;
    int ptab_colbits = 32;
    int64_t npar = 0;            // parent sector size
    // reference labels and normalizations (host, computed on demand)
    std::vector<uint64_t> labels;
    std::vector<double> ref_norm;
    bool labels_ready = false;
    // split-kernel tables (built once, on first use)
    mutable std::once_flag split_once;
    mutable std::unique_ptr<SplitBasis> split;
    // SpinBasisGeneral: the blocks in stage order, and the group on the
    // device: per element four 256-entry byte tables of the permutation, its
    // xor constant and its character (interleaved complex)
    std::vector<GenMap> gmaps;
    DevBuf<uint32_t> gtab, gc;
    DevBuf<double> gchi;
    // wide bases (general maps on more than 32 sites, up to 256): the orbit
    // minima as four words each, ascending; the group as a site table per
    // element (byte e * L + i = target of site i) and its xor constants
    // (four words per element); characters in gchi
    int wide = 0;
    DevBuf<uint64_t> wreps;
    DevBuf<uint8_t> wperm;
    DevBuf<uint64_t> wc;
    std::vector<WState> wlabels;   // reference labels (host, on demand)
};

struct OpHost {
    const BasisHost *basis = nullptr;
    int is_complex = 0;          // any complex matrix element
    int hermitian = 1;
    // action terms (P'HP semantics, group averaged), adjoint form for the pull kernel
    std::vector<CTerm> diag_terms;
    std::vector<CTerm> off_terms;
    std::vector<GroupOfTerms> groups;
    DevBuf<TermDev> d_diag, d_off;
    DevBuf<GroupOfTerms> d_groups;
    // adjoint operator (for transposed / adjoint products)
    std::vector<CTerm> diag_terms_adj, off_terms_adj;
    std::vector<GroupOfTerms> groups_adj;
    DevBuf<TermDev> d_diag_adj, d_off_adj;
    DevBuf<GroupOfTerms> d_groups_adj;
    // fast forms (hop amplitude tables, diagonal classes) of both term lists
    FastTerms fast, fast_adj;
    DevBuf<HopDev> d_hops, d_hops_adj;
    DevBuf<DiagClass> d_dcls, d_dcls_adj;
    DevBuf<TermDev> d_drest, d_drest_adj;
    // assembled sector matrix (qs_op_assemble), used for plain products;
    // want_sell defers the assembly to the first plain product (its stream)
    std::unique_ptr<SellHost> sell;
    mutable bool want_sell = false;
    mutable std::mutex sell_mutex;
    // hop groups classified for the split kernel (built once, on first use)
    mutable std::once_flag split_once;
    mutable SplitHops split_hops, split_hops_adj;
    // and for the passes of its tiled form, per (kA, kB)
    mutable std::mutex pass_mutex;
    mutable std::vector<std::unique_ptr<SplitPassHops>> pass_hops;
    // wide bases: the term lists (plain products apply the adjoint
    // couplings, as above), grouped by flip mask
    std::vector<WTerm> wdiag, woff, wdiag_adj, woff_adj;
    std::vector<WGroupDev> wgroups, wgroups_adj;
    DevBuf<WTermDev> d_wdiag, d_woff, d_wdiag_adj, d_woff_adj;
    DevBuf<WGroupDev> d_wgroups, d_wgroups_adj;
    // user diagonal
    std::vector<cd> user_diag;
    bool has_user_diag = false;
    DevBuf<double> d_user_diag;      // interleaved complex
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int lanes_per_row = 0;       // matvec kernel variant (0 = default heuristic)
    int graphs = 0;              // replay the expm recursion's factorizations from a HIP graph
    // the hop kernel's pre-scaled x (parent-table sectors), kept between
    // products: one buffer per stream, so products of concurrent jobs on
    // different streams never share one (each stream orders its own writes
    // and reads).  A buffer only grows; growing it frees the old one with
    // hipFree, which waits for the products still reading it.
    struct XsBuf {
        hipStream_t s;
        DevBuf<char> buf;
    };
    mutable std::mutex xs_mutex;
    mutable std::vector<std::unique_ptr<XsBuf>> xs_bufs;
    // the pre-scale buffer of stream s, at least `bytes` long
    char *xs_for(hipStream_t s, size_t bytes) const;
    // staging buffers for host vectors
    DevBuf<char> stage_x, stage_y;
    // basis without symmetry: the extracted diagonal alone (structurally
    // diagonal terms + user diagonal), which a conjugated product scales by
    // alpha while the other terms take conj(alpha)
    std::unique_ptr<OpHost> extracted;
    ~OpHost();
};

GroupDev group_to_dev(const GroupHost &h);

// Stream the solvers launch on for `op` from the calling thread: a worker
// stream installed with StreamScope (the serving loop runs several jobs on
// one operator at once), else the operator's own stream.
hipStream_t op_stream(const OpHost &op);
struct StreamScope {
    explicit StreamScope(hipStream_t s);
    ~StreamScope();
    hipStream_t prev;
};

// ---------------------------------------------------------- host logic ---
uint64_t full_mask(int L);
uint64_t bit_reverse(uint64_t u, int L);
uint64_t rotl(uint64_t u, int k, int L);
uint64_t apply_point(const GroupHost &g, int p, uint64_t u);
int64_t binom(int n, int k);

void build_group(BasisHost &b);   // throws Error(QS_EARG) for invalid blocks
void build_group_general(BasisHost &b);   // from b.gmaps (SpinBasisGeneral)
// element e of the group (e = t * npoint + p, or the general list index)
uint64_t group_apply(const GroupHost &g, int e, uint64_t u);
cd group_char(const GroupHost &g, int e);
void build_basis_device(BasisHost &b);
void compute_reference_labels(BasisHost &b);
// wide bases (qs_wide.hip, qs_wterms.cpp, qs_labels.cpp)
void build_group_wide(BasisHost &b);            // the group of b.gmaps on up to 256 sites
void build_wide_basis(BasisHost &b);            // group from b.gmaps, orbit minima on the device
void wide_enumerate(BasisHost &b);              // device enumeration of the orbit minima
void compute_wide_labels(BasisHost &b);         // reference labels and normalizations
WState wide_group_apply(const GroupHost &g, int e, const WState &u);
void wide_local(char op, int bit, int pauli, bool &alive, int &out, cd &amp);
void launch_matvec_wide(const OpHost &op, int dtype, const void *x, void *y, double are, double aim, double bre,
                        double bim, bool adjoint, bool conj_x, bool conj_y, hipStream_t s, double *dot_partial);

struct Coupling {
    std::string op;
    std::vector<int> sites;   // 1-based
    cd coef;
};
// wide-basis term lists of an operator (qs_wterms.cpp); check_symm: reject a
// list the group does not leave invariant
void setup_wide_terms(OpHost &o, const std::vector<Coupling> &cs, bool check_symm);
std::vector<CTerm> compile_couplings(const std::vector<Coupling> &cs, int L, int pauli);
std::vector<CTerm> consolidate_cterms(const std::vector<CTerm> &t);
std::vector<CTerm> symmetrize_cterms(const std::vector<CTerm> &t, const BasisHost &b);
void split_terms(const std::vector<CTerm> &all, std::vector<CTerm> &diag, std::vector<CTerm> &off,
                 std::vector<GroupOfTerms> &groups);

// small dense FP64 linear algebra (host)
void sym_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V);
void herm_eig(int n, std::vector<cd> &A, std::vector<double> &w, std::vector<cd> &V);
void expm_dense(int n, const std::vector<cd> &A, std::vector<cd> &E);

// --------------------------------------------------------- kernels -------
struct VecWork;
// y = beta*y + alpha*K x on the sector; K = M (adjoint=false) or M^dagger.
// conj_x / conj_y conjugate x on load and y on load and store.
// dot_partial (optional): matvec_grid(Ns) x 2 doubles receiving per-block
// FP64 partial sums of <x, y_new> computed before y is rounded to its dtype.
// xs_scratch (optional): device memory of at least 16 Ns bytes for the
// pre-scaled x of parent-table sectors; with it the product uses only
// kernels whose tables qs_op_create built and allocates nothing.
void launch_matvec(const OpHost &op, int dtype, const void *x, void *y, double are, double aim,
                   double bre, double bim, bool adjoint, bool conj_x, bool conj_y, hipStream_t s,
                   double *dot_partial = nullptr, void *xs_scratch = nullptr);
// Device address of the fixed scratch array of the no-allocation product on
// the current device (looked up once per device)
void *noalloc_scratch();
int matvec_grid(int64_t n);
// the two-bit hop kernel applies to this operator (plain or adjoint list)
bool hop_eligible(const OpHost &op, bool adjoint);
// the split kernel applies (hop couplings, implicit basis with a particle number)
bool split_eligible(const OpHost &op, bool adjoint);
bool split_lean_ok(const OpHost &op, bool adjoint);
// split variants: QS_VARIANT_SPLIT + k, k < QS_SPLIT_LEAN single pass
// (k fewer resident workgroups), k = QS_SPLIT_LEAN the lean kernel in one
// pass, k = QS_SPLIT_TILED + d the lean kernel's tiled two-pass form with
// tile budget 2^(20 + d) bytes
constexpr int QS_SPLIT_LEAN = 3;
constexpr int QS_SPLIT_TILED = 4;
constexpr int QS_VARIANT_SPLIT_MAX = QS_VARIANT_SPLIT + QS_SPLIT_TILED + 3;
// Assemble the sector matrix of the plain product (hop-eligible operators),
// and read it back as CSR with duplicate columns merged, in ascending column
// order (rowptr n+1, col / val nnz; val interleaved complex).
void sell_assemble(OpHost &op, hipStream_t s);
void sell_to_csr(const OpHost &op, std::vector<int64_t> &rowptr, std::vector<int64_t> &col,
                 std::vector<double> &val);
// <x, A x> in FP64 with x stored in `dtype` (fused into the matvec epilogue); y is scratch
cd rayleigh_numerator(OpHost &op, int dtype, const void *x, void *y, VecWork &wk);

// ------------------------------------------------------ batched launches --
// One launch carrying several jobs: each job owns a contiguous range of the
// grid's blocks and sees exactly the block partition of its own single
// launch, so its results are bitwise those of the single launch.
enum VecOp { VOP_GRAM = 0, VOP_REDUCE, VOP_UPDATE, VOP_NORM2, VOP_AXPBY, VOP_COMBINE };
// Axpby coefficient a taken from device memory c (a reduction result):
// CO_NEG a = -(c0 + i c1), CO_NEG_RE a = -c0, CO_NEG_SQRT a = -sqrt(c0),
// CO_INV_SQRT a = 1 / sqrt(c0); CO_HOST: a is the job's (ar, ai).
enum CoefMode { CO_HOST = 0, CO_NEG, CO_NEG_RE, CO_NEG_SQRT, CO_INV_SQRT };
struct VecJob {
    const void *V;      // basis block (column-major, ld n)
    const void *x;      // input vector
    void *w;            // in/out vector (update, axpby, combine)
    double *partial;    // per-block partial sums
    double *aux;        // reduce: results; update: coefficients h; combine: C
    int64_t n;
    int m, nb, b0, k, b_zero;
    double ar, ai, br, bi;
    const double *acoef;   // axpby: device coefficient source (CoefMode cmode)
    int cmode;
};
constexpr int kMaxVecJobs = 24;
struct VecJobs {
    int njobs;
    VecJob job[kMaxVecJobs];
};
// Launch `op` over jobs (b0 and the grid are filled in here).  Up to
// kMaxVecJobs jobs travel in the kernel arguments; more go through `table`
// (one launch, each job on max(blocks) blocks) or, without one, to further
// launches on the same stream.  A job's blocks do the same work either way.
void launch_vec_jobs(int dtype, int op, std::vector<VecJob> &jobs, const std::vector<int> &blocks, hipStream_t s,
                     JobTable *table = nullptr);
int ew_blocks(int64_t n);
// w_j -= V_j h_j then partials of V_j^H w_j and |w_j|^2 (m + 1 per block, as
// k_update_gram) for jobs sharing one column count m (job.aux = h); false
// (nothing launched) when m is too large for the fused kernel
bool launch_updgram_jobs(int dtype, std::vector<VecJob> &jobs, hipStream_t s, JobTable *table);

// Krylov steps of a small assembled sector in one single-workgroup launch
// (qs_vec.hip): SMALL_DGKS runs steps j0 .. steps-1 of a thick-restart
// Lanczos cycle (per-step scalars `per` apart, as eigsh_device lays them
// out); SMALL_LANCZOS / SMALL_ARNOLDI a whole expm factorization from vin
// (FactScalars layout).  krylov_small_ok: the operator's plain products read
// stored rows, n <= kSmallMaxN and at most kSmallMaxCols basis columns.
enum SmallMode { SMALL_DGKS = 0, SMALL_LANCZOS = 1, SMALL_ARNOLDI = 2 };
// above ~2k rows one CU's L2 bandwidth makes the launch slower than the
// per-operation launches over the whole chip (measured: 4.6k rows 32 vs 13 ms
// per evolve)
constexpr int64_t kSmallMaxN = 2048;
constexpr int kSmallMaxCols = 128;
bool krylov_small_ok(const OpHost &op, int cols);
void krylov_small(const OpHost &op, int dtype, int mode, int steps, int j0, void *V, const void *vin, double *dd,
                  int per, hipStream_t s);
// assemble an operator marked for assembly (QS_VARIANT_SELL) now, on stream s
void ensure_assembled(const OpHost &op, hipStream_t s);

// Batched matvec: y_j = M_j x_j for several operators in one launch (all
// must take the hop kernel with one wave per row tile).
struct MatvecJob {
    int op;            // index into the batch's operator list
    const void *x;
    void *y;
    // assembled operators only: y += a z with a from device memory
    // (CoefMode zmode of zc), then per-block FP64 partials of <x, y> to
    // dot[2 b] (b < ceil(N / 256)); the next step of a Lanczos recurrence
    const void *z = nullptr;
    const double *zc = nullptr;
    int zmode = 0;
    double *dot = nullptr;
};
struct BatchMatvec;   // device copies of the operators' kernel arguments
std::unique_ptr<BatchMatvec, void (*)(BatchMatvec *)> make_batch_matvec(const std::vector<const OpHost *> &ops,
                                                                         int dtype);
void launch_batch_matvec(BatchMatvec &bm, const std::vector<MatvecJob> &jobs, hipStream_t s);
// One Lanczos product of an assembled operator (the MatvecJob fields z, zc,
// zmode, dot for a single product, arguments by value); false when the
// operator's plain products do not read stored rows
bool matvec_lanczos(const OpHost &op, int dtype, const void *x, void *y, const void *z, const double *zc, int zmode,
                    double *dot, hipStream_t s);
// out[0..1] = the sum of nb interleaved complex partials
void vec_reduce_dev(const double *partial, int nb, double *out, hipStream_t s);
bool batch_matvec_ok(const std::vector<const OpHost *> &ops, int dtype);

// Sector <-> parent space (qs_project.hip).  kind 0: the parent vector is
// indexed like the basis' own parent (colex rank, or the state without nup);
// kind 1: by the state (2^L entries).  out = P' parent;  parent (+)= P x for
// ncols columns (column-major, leading dimensions parent_length and Ns).
int64_t parent_length(const BasisHost &b, int kind);
void basis_project(const BasisHost &b, int dtype, int kind, const void *parent, void *out, hipStream_t s);
void basis_expand(const BasisHost &b, int dtype, int kind, const void *x, void *parent, int ncols, bool accumulate,
                  hipStream_t s);

// _krylov_expmv_times of many (operator, v, time grid) tasks in lockstep
// (qs_batch.cpp); out: Ns x scales.size() of the complex dtype per task.
struct ExpmTask {
    int op;
    const void *v;
    std::vector<cd> scales;
    void *out;
};
void expm_times_batch(const std::vector<const OpHost *> &ops, int dtype, const std::vector<ExpmTask> &tasks,
                      double tol, int krylov_dim, hipStream_t s, std::vector<int> &nmatvec);

// eigsh / _krylov_expmv_times of one operator on stream s with host
// vectors (the bodies of qs_eigsh and qs_expm_multiply; the serving loop's
// workers call them with their own streams).
void eigsh_host(OpHost &o, hipStream_t s, int dtype, int k, int which, double tol, int maxiter, int ncv,
                const void *v0, double *vals, void *vecs, int *nconv, int *n_matvec);
void expm_host(OpHost &o, hipStream_t s, int dtype, const void *v, int64_t n_times, const double *scales, double tol,
               int krylov_dim, void *out, int *n_matvec);

// evolve with dynamic terms (qs_rk45.hip): the reference's adaptive RK45
void evolve_dynamic_host(const OpHost &h0, const std::vector<const OpHost *> &dyn, qs_drive_fn drive, void *user,
                         int dtype, const void *v0, int64_t ncols, double t0, const std::vector<double> &times,
                         int imag_time, double rtol, double atol, double max_step, void *out, int64_t *n_steps,
                         int64_t *n_matvec);

// vector kernels (qs_vec.hip); all reductions are deterministic (fixed order)
struct VecWork {
    explicit VecWork(hipStream_t st) : s(st) {}
    hipStream_t s;            // the stream every launch using this workspace runs on
    SBuf<double> partial;     // per-block partial sums
    SBuf<double> result;      // reduced values (device)
    std::vector<double> host;
};
int vec_blocks(int64_t n);
// h[j] = <V[:, j], w> for j < m (complex: conj(V) . w); V column-major with ld = n.
void gram_dots(int dtype, const void *V, int64_t n, int m, const void *w, VecWork &wk,
               hipStream_t s, std::vector<cd> &h);
// w -= V[:, :m] h  and return ||w||^2
double gram_update(int dtype, const void *V, int64_t n, int m, const std::vector<cd> &h, void *w,
                   VecWork &wk, hipStream_t s);
double vec_norm2(int dtype, const void *x, int64_t n, VecWork &wk, hipStream_t s);
// Device-scalar forms: results stay in device memory (out / nrm2_out, two
// doubles per value), coefficients are read from there; nothing waits for
// the host.  vec_reserve sizes wk for up to mmax Gram columns beforehand
// (the workspace must not be reallocated while launches are queued).
void vec_reserve(VecWork &wk, int64_t n, int mmax);
// With dgks_h1 / dgks_n2 (the first pass's h and |w'|^2, device) these are
// the conditional second pass of ARPACK's DGKS step: skipped on the device
// (zeros for h, |w'|^2 copied for the norm) when |w'| > 0.717 |w|.
void gram_dots_dev(int dtype, const void *V, int64_t n, int m, const void *w, VecWork &wk, double *out,
                   hipStream_t s, const double *dgks_h1 = nullptr, const double *dgks_n2 = nullptr);
void gram_update_dev(int dtype, const void *V, int64_t n, int m, const double *h, void *w, VecWork &wk,
                     double *nrm2_out, hipStream_t s, const double *dgks_h1 = nullptr,
                     const double *dgks_n2 = nullptr);
// ARPACK's DGKS step (dsaitr / znaitr) on w against V[:, :m]: the first
// Gram pass (h1), w' = w - V h1 with |w'|^2 (n2a), and, only if |w'| <=
// 0.717 |w| (decided on the device), the second pass (h2, else zeros) and
// w'' = w' - V h2; n2 = the final |w|^2.  The first pass's update and the
// second pass's Gram product share one read of V.  Results in device memory.
void gram_dgks_dev(int dtype, const void *V, int64_t n, int m, void *w, VecWork &wk, double *h1, double *h2,
                   double *n2a, double *n2, hipStream_t s);
// block Gram product G = V^H W (V: n x m, W: n x p, ld n) on the matrix
// cores, FP64 accumulation; out: m * p interleaved complex, column-major
void gram_block_dev(int dtype, const void *V, int64_t n, int m, const void *W, int p, VecWork &wk, double *out,
                    hipStream_t s);
void gram_block(int dtype, const void *V, int64_t n, int m, const void *W, int p, VecWork &wk, hipStream_t s,
                std::vector<cd> &G);
void vec_norm2_dev(int dtype, const void *x, int64_t n, VecWork &wk, double *out, hipStream_t s);
void vec_axpby_dev(int dtype, const double *acoef, int cmode, const void *x, cd b, void *y, int64_t n,
                   hipStream_t s);
void vec_scale_copy(int dtype, const void *x, void *y, int64_t n, cd scale, hipStream_t s);
void vec_axpby(int dtype, cd a, const void *x, cd b, void *y, int64_t n, hipStream_t s);
// Y[:, :k] = V[:, :m] C  (C is m x k column-major, complex or real)
void basis_combine(int dtype, const void *V, int64_t n, int m, const std::vector<cd> &C, int k,
                   void *Y, hipStream_t s);
void vec_random(int dtype, void *y, int64_t n, uint64_t seed, hipStream_t s);
// Y = V C on the matrix cores when there are at least kMfmaMinCols output
// columns and the storage is FP64 (real or complex); false (nothing
// launched) otherwise.  C is the device copy, interleaved complex m x k.
constexpr int kMfmaMinCols = 16;
bool combine_on_mfma(int dtype, const void *V, int64_t n, int m, const double *C, int k, void *Y, hipStream_t s);
void vec_convert(int from_dtype, const void *x, int to_dtype, void *y, int64_t n, hipStream_t s);
size_t dtype_size(int dtype);
bool dtype_complex(int dtype);

// Host side of an expm Krylov factorization (operators.jl:1655-1756): the
// (steps+1) x steps projected matrix and the quantities the reference's
// a-posteriori estimate needs.  Shared by the single and batched drivers so
// both do the same host arithmetic.
// Device layout of one factorization's scalars (doubles): |v|^2, then per
// column the two Gram passes' coefficients (steps + 1 complex each) and the
// squared norm of the new residual.
struct FactScalars {
    int steps;
    explicit FactScalars(int st) : steps(st) {}
    size_t per() const { return (size_t)4 * (steps + 1) + 2; }
    size_t beta2() const { return 0; }
    size_t h1(int col) const { return 2 + (size_t)col * per(); }
    size_t h2(int col) const { return h1(col) + (size_t)2 * (steps + 1); }
    size_t nrm2(int col) const { return h2(col) + (size_t)2 * (steps + 1); }
    size_t total() const { return 2 + (size_t)steps * per(); }
};

struct KrylovProj {
    std::vector<cd> H;   // row-major (steps+1) x steps
    int steps = 0, completed = 0;
    double beta = 0, terminal = 0;
    int64_t dim = 0;
    bool hermitian = true;
    cd h(int i, int j) const { return H[(size_t)i * steps + j]; }
    void lanczos_entry(int col, cd alpha, double tn);
    void arnoldi_column(int col, const std::vector<cd> &h1, const std::vector<cd> &h2, double tn);
    // records tn; true (and completed = col + 1) on breakdown; f = squared
    // Frobenius norm of H's entries so far
    bool breakdown(int col, double tn, double f);
    // the per-step bookkeeping of a factorization queued with its scalars on
    // the device (beta, H entries, breakdown test in step order); returns the
    // number of steps the reference performs (= matvecs)
    int replay(const FactScalars &L, const std::vector<double> &d);
    // beta exp(scale H_c)[:, 0] and the error estimate (_krylov_project)
    double expm_column(cd scale, std::vector<cd> &coef) const;
    // Hermitian: coefficients for every scale from one eigen-decomposition
    void lanczos_times(const std::vector<cd> &scales, std::vector<cd> &coefs, std::vector<double> &ests) const;
};

// lanczos_full / lanczos_iter recurrence (qs_krylov.cpp): Q gets the used
// rows (Ns x m column-major); alphas / betas the SymTridiagonal entries.
void lanczos_device(OpHost &op, int dtype, const void *v0, int m, bool full_ortho, double eps, void *Q,
                    std::vector<double> &alphas, std::vector<double> &betas, int &used);

// Krylov solvers (qs_krylov.cpp); vectors are device buffers of `dtype`
// expm factorizations use three-term Lanczos (the reference's Hermitian
// path) only for Hermitian operators in FP64 storage
bool lanczos_ok(const OpHost &op, int dtype);
struct EigshResult {
    std::vector<double> vals;    // k (re) or 2k (re, im)
    int nconv = 0, nmatvec = 0;
};
// non-Hermitian operators: Krylov-Schur Arnoldi (qs_arnoldi.cpp); complex
// dtypes only, vals = 2k interleaved (re, im)
void eigs_device(OpHost &op, int dtype, int k, int which, double tol, int maxiter, int ncv, const void *v0_dev,
                 EigshResult &res, void *vecs_dev);
// eigenvalues w and right eigenvectors V (row-major, column i for w[i]) of a
// general row-major n x n complex matrix
void gen_eig(int n, const std::vector<cd> &A, std::vector<cd> &w, std::vector<cd> &V);
void eigsh_device(OpHost &op, int dtype, int k, int which, double tol, int maxiter, int ncv, const void *v0_dev,
                  EigshResult &res, void *vecs_dev);
// out_dev: Ns x n_times (column-major) of the complex dtype
void expm_times_device(OpHost &op, int dtype, const void *v_dev, const std::vector<cd> &scales, double tol,
                       int krylov_dim, void *out_dev, int &nmatvec);

}  // namespace qs
