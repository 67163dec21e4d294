/*
 * qsgpu.h -- C ABI of libqsgpu.so, the MI355X (gfx950) implementation of
 * QuSpin.jl's spin-1/2 matrix-free Hamiltonian action and the Krylov solvers
 * that consume it.
 *
 * The reference has no FFI layer: its boundary is Julia's AbstractMatrix
 * protocol (size / eltype / ishermitian / mul!) that QuantumLinearOperator
 * implements (operators.jl:3128-3245) and that Arpack.eigs (operators.jl:3384)
 * and the Krylov exponential (operators.jl:1642-1653) consume.  Each entry
 * point below names the reference function it stands in for; the Julia
 * `ccall` shim in quspin.jl_amd/julia/QuSpinGPU.jl binds them.
 *
 * Conventions
 *  - Every function returns a status code (QS_OK on success).  The message of
 *    the last failure on the calling thread is read with qs_last_error().
 *  - Handles are opaque and owned by the library; release them with the
 *    matching *_destroy.  Calls on one handle are serialised by the caller;
 *    calls on different handles may run concurrently.
 *  - Sites are 1-based, as in QuSpin.jl.  Complex numbers are interleaved
 *    (re, im) doubles.  Vectors are contiguous; a block of ncols vectors is
 *    column-major with leading dimension Ns.
 *  - `on_device` = 1 means the vector pointers are HIP device pointers that
 *    the kernels read directly; 0 means host pointers that are staged.
 */
#ifndef QSGPU_H
#define QSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the Julia shim maps them to the reference's exceptions. */
#define QS_OK 0
#define QS_EARG 1     /* ArgumentError      (e.g. basis.jl:655,686-687,717-723) */
#define QS_EDIM 2     /* DimensionMismatch  (operators.jl:3154-3157) */
#define QS_ENOCONV 3  /* ErrorException: Krylov did not converge (operators.jl:1782-1784) */
#define QS_EOOM 4     /* device or host allocation failed */
#define QS_EDEVICE 5  /* HIP runtime error */

/* A symmetry block argument that is `nothing` in Julia. */
#define QS_UNSET (-1000000)

/* Element types of vectors crossing the boundary. */
#define QS_F64 0   /* Float64 */
#define QS_C128 1  /* ComplexF64 */
#define QS_F32 2   /* Float32 storage (FP32-stream form) */
#define QS_C64 3   /* ComplexF32 storage (FP32-stream form) */

/* eigsh `which` (operators.jl:3375-3377). */
#define QS_WHICH_SA 0
#define QS_WHICH_LA 1
#define QS_WHICH_SM 2
#define QS_WHICH_LM 3
#define QS_WHICH_BE 4

/* qs_op_create flags (operators.jl:2492-2499 keyword arguments). */
#define QS_CHECK_SYMM 1
#define QS_CHECK_HERM 2
#define QS_CHECK_PCON 4
/* Complex element type even when every matrix element is real: the
 * caller's coefficients or `diagonal` are complex-typed, which makes the
 * reference's element type T complex (_coefficient_type, operators.jl:329-341;
 * promote_type with eltype(diagonal), operators.jl:2523-2527), so products
 * and solvers take complex vectors. */
#define QS_FORCE_COMPLEX 8
/* qs_basis_create_general flags: the wide (four-word, stored) form at any L */
#define QS_BASIS_EXPLICIT 1

typedef struct qs_basis qs_basis;
typedef struct qs_op qs_op;

/* Last error message of the calling thread; returns its full length. */
size_t qs_last_error(char *buf, size_t len);

/* Library / device information: HIP device name and total HBM bytes. */
int qs_device_info(char *name, size_t len, int64_t *hbm_bytes, int *n_cu);

/* ---------------------------------------------------------------- basis --
 * SpinBasis1D(L; nup, pauli, a, kblock, pblock, zblock, pzblock, zAblock,
 * zBblock)  -- basis.jl:622-777.  nup = QS_UNSET enumerates every state.
 * The sector is enumerated on the GPU (orbit-minimum canonical states in
 * ascending order, which is the reference's column order).  flags: 0. */
int qs_basis_create(int L, int nup, int pauli, int a, int kblock, int pblock, int zblock,
                    int pzblock, int zAblock, int zBblock, int flags, qs_basis **out);
/* SpinBasisGeneral(L; Nup, pauli, block_order, blocks...) -- basis.jl:530-620
 * (spin 1/2): the nup sector (QS_UNSET: every state) reduced by nblocks site
 * maps in stage order (the reference's block_order).  maps: nblocks x L
 * entries, each row one map in QuSpin's zero-based labels (any entry zero or
 * negative; -v inverts the site and sends it to site v) or Julia's one-based
 * labels (negative = inverted), as _normalize_general_spin_map reads them
 * (basis.jl:282-304); qnumbers: each block's quantum number q, eigenvalue
 * cis(2 pi mod(q, order) / order) with the map's order (basis.jl:306-323).
 * L <= 256; the maps may generate at most 256 elements.  L <= 32: sectors
 * take two-bit hop couplings (the parent-table kernels).  L > 32 with
 * blocks, L > 63, or flags & QS_BASIS_EXPLICIT: a wide basis (basis.jl:
 * 361-445; the reference's WideSpinBasis above 63 sites), which needs nup:
 * its orbit minima are stored as four 64-bit words each (site i = bit i % 64
 * of word i / 64) and products take any monomial couplings (one kernel,
 * group orbit and binary search per neighbour).  Errors: the reference's
 * ArgumentError messages for a map that is not a permutation, a flip that
 * leaves the particle sector, an empty sector, a missing nup above 32 sites. */
int qs_basis_create_general(int L, int nup, int pauli, int nblocks, const int32_t *maps, const int32_t *qnumbers,
                            int flags, qs_basis **out);
int qs_basis_destroy(qs_basis *b);

/* length(basis) / basis.Ns -- basis.jl:779 */
int qs_basis_size(const qs_basis *b, int64_t *ns);

/* states(basis) -- basis.jl:786: the reference's representative label per
 * column (symmetry_basis.jl:215-232), in column order. */
int qs_basis_states(const qs_basis *b, uint64_t *out);

/* Orbit-minimum canonical state per column (the key the GPU kernels index
 * by; equals states(basis) wherever the reference label is the orbit min). */
int qs_basis_reps(const qs_basis *b, uint64_t *out);

/* states(basis) (which = 0) or the orbit minima (which = 1) as four 64-bit
 * words per column (site i = bit i % 64 of word i / 64), for any basis; the
 * only form that holds the states of a basis of more than 64 sites
 * (basis.jl:802 states(::WideSpinBasis), UInt256 values).  out: 4 Ns words.
 * qs_basis_states / qs_basis_reps refuse such a basis. */
int qs_basis_states_wide(const qs_basis *b, int which, uint64_t *out);

/* qs_basis_lookup for states given as four 64-bit words each (any basis; the
 * only form for bases of more than 64 sites): column of each state's orbit
 * (-1 outside the sector) and its amplitude there -- state_index /
 * get_amp of a WideSpinBasis (basis.jl:1049-1075). */
int qs_basis_lookup_wide(const qs_basis *b, const uint64_t *states, int64_t n, int64_t *index, double *amplitude);

/* Per column: number of distinct parent states in the orbit, and the
 * reference's normalization(basis, states(basis)[i]) (discrete_basis.jl:2401-2408). */
int qs_basis_norms(const qs_basis *b, uint32_t *orbit_size, double *ref_norm);

/* _basis_requires_complex (symmetry_basis.jl:624-629). */
int qs_basis_is_complex(const qs_basis *b, int *is_complex);

/* state_index(basis, s) -- basis.jl:1016-1027 on the column keys: index of
 * the column whose orbit contains states[i] (0-based, -1 if none), and the
 * amplitude <states[i]|column> (interleaved complex, may be NULL). */
int qs_basis_lookup(const qs_basis *b, const uint64_t *states, int64_t n, int64_t *index,
                    double *amplitude);

/* ------------------------------------------------------------- operator --
 * QuantumLinearOperator(basis, terms; diagonal, check_symm, check_herm,
 * check_pcon) -- operators.jl:2492-2552.
 *   opstr[t]        operator string of term t ("+-", "zz", "z", "x", "y", "I")
 *   n_couplings[t]  couplings of term t
 *   coeff           2 doubles (re, im) per coupling, terms in order
 *   sites           len(opstr[t]) 1-based sites per coupling, terms in order
 *   diag            NULL or 2*Ns doubles: the user `diagonal` keyword
 * The operator realises the reference's P'·H·P semantics in the sector. */
int qs_op_create(const qs_basis *b, int n_terms, const char *const *opstr,
                 const int32_t *n_couplings, const double *coeff, const int32_t *sites,
                 const double *diag, int flags, qs_op **out);
int qs_op_destroy(qs_op *op);

/* eltype / ishermitian / size -- operators.jl:3228-3245. */
int qs_op_info(const qs_op *op, int *is_complex, int *is_hermitian, int64_t *ns);

/* set_diagonal!(op, d) -- operators.jl:3301-3309 (2*Ns doubles). */
int qs_op_set_diagonal(qs_op *op, const double *diag);

/* Stream the operator launches on (hipStream_t); NULL selects the
 * library's own stream.  qs_op_stream returns the one in use. */
int qs_op_set_stream(qs_op *op, void *stream);
int qs_op_stream(const qs_op *op, void **stream);

/* Matvec kernel variant: 0 = automatic; 1, 2, 4 or 8 = the general kernel
 * with that many lanes cooperating on one row; QS_VARIANT_HOP + k (k = 0, 1,
 * 2) = the two-bit hop kernel (couplings that flip two sites and test only
 * those two, L <= 32; symmetry sectors also need the parent lookup table)
 * with 2^k wavefronts sharing each row tile's couplings; QS_VARIANT_HOP + 3 =
 * the hop kernel with one wavefront per tile that, in a symmetry sector,
 * first writes x / sqrt(norm) to an FP64 buffer held by the operator and then
 * gathers one value per neighbour.  No reference counterpart: the reference
 * has one CPU loop. */
#define QS_VARIANT_HOP 16
/* The split kernel: two-bit hop couplings on a basis without symmetry or
 * with spin inversion only, with a particle number, L <= 32.  Rows sharing
 * their high bits form a contiguous block, so hops among the high bits read
 * whole neighbour blocks and hops among the low bits stay in LDS;
 * QS_VARIANT_SPLIT + k (k = 0, 1, 2) is one pass that keeps k fewer
 * workgroups per CU resident than fit; QS_VARIANT_SPLIT + 3 is the lean
 * form (neighbour rows from per-class tables, high hops in scalar
 * registers); QS_VARIANT_SPLIT + 4 + d (d = 0..3) is the lean form in two
 * passes, each over tiles of blocks whose hops stay inside the tile (at
 * most 2^(20+d) bytes of x per tile, one tile per XCD at a time), the second
 * adding to the first's y.  No reference counterpart. */
#define QS_VARIANT_SPLIT 32
/* Plain products from the assembled sector matrix (qs_op_assemble).  The
 * reference's counterpart is the assembled Hamiltonian, static_fmt = :csr. */
#define QS_VARIANT_SELL 48
int qs_op_set_variant(qs_op *op, int lanes_per_row);

/* Opt in to HIP-graph replay of the Krylov exponential's halving
 * recursion: on a non-blocking stream (qs_op_set_stream) its second and later
 * factorizations replay one captured graph (every step queued with device
 * scalars, operators.jl:1655-1731).  While a capture is open the HIP runtime
 * refuses work on the legacy default stream from every thread of the
 * process, so enable it only where no other thread uses that stream during
 * the solve.  Off by default.  No reference counterpart. */
int qs_op_set_graphs(qs_op *op, int enable);

/* Hamiltonian(...; static_fmt=:csr) assembly -- operators.jl:461-576
 * (_assemble: P'.H_par.P for a sector, sparse_formats.jl:21-50 storage):
 * builds the sector matrix of the plain product on the device (sliced ELL,
 * 64-row slices) from the same neighbour enumeration as the matrix-free
 * kernel.  Afterwards plain products read the stored rows (variant 0 or
 * QS_VARIANT_SELL); transposed / adjoint products stay matrix-free.  Needs
 * two-bit hop couplings (QS_VARIANT_HOP's condition). */
int qs_op_assemble(qs_op *op);

/* The assembled matrix as CSR (SparseMatrixCSR, sparse_formats.jl:21-50):
 * duplicate columns merged and exact zeros dropped (operators.jl:567-572),
 * columns ascending; val interleaved complex (2*nnz doubles).  With rowptr,
 * col and val all NULL only *nnz is written; otherwise *nnz must hold the
 * capacity on entry. */
int qs_op_csr(const qs_op *op, int64_t *nnz, int64_t *rowptr, int64_t *col, double *val);

/* Autotuner: times every variant for `dtype` on device-resident random
 * vectors (`repeats` launches each), keeps the fastest on the operator and
 * returns it with its time.  No reference counterpart. */
int qs_op_tune(qs_op *op, int dtype, int repeats, int *lanes_per_row, double *best_ms);

/* mul!(Y, op, X, alpha, beta) for ncols column vectors --
 * operators.jl:3147-3208; transposed/conjugated select transpose(op),
 * conj(op) and adjoint(op) (operators.jl:3286-3296).
 * dtype: QS_F64 / QS_C128 (FP64 form) or QS_F32 / QS_C64 (FP32 stream,
 * FP64 row accumulation).  Real dtypes need a real operator. */
int qs_op_matvec(qs_op *op, int dtype, const void *x, void *y, int64_t ncols, double alpha_re,
                 double alpha_im, double beta_re, double beta_im, int transposed, int conjugated,
                 int on_device);

/* The same product on device vectors with no device allocation: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It runs
 * only kernels whose tables qs_op_create built (the two-bit hop kernel or the
 * general kernels; never the lazily assembled or split forms), and in a
 * parent-table sector (translations / parity with a particle number) the
 * pre-scaled x goes to a fixed scratch array of QS_NOALLOC_SCRATCH_BYTES per
 * GPU (16 bytes per complex, 8 per real row: at most 2^21 complex or 2^22 real
 * rows; larger sectors are rejected with QS_EARG).  x and y are device
 * pointers and must not alias (the copy the reference makes of an aliased
 * input, operators.jl:3158, needs a buffer).  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (concurrent streams or
 * threads) race and give wrong results; serialise them. */
#define QS_NOALLOC_SCRATCH_BYTES (32u << 20)
int qs_op_matvec_noalloc(qs_op *op, int dtype, const void *x, void *y, int64_t ncols, double alpha_re,
                         double alpha_im, double beta_re, double beta_im, int transposed, int conjugated);

/* eigsh(op; k, which, tol, maxiter, ncv, v0, return_eigenvectors) --
 * operators.jl:3367-3408 (ARPACK "SA"/"LA"/... semantics).  Hermitian
 * operators: thick-restart Lanczos with full FP64-accumulated
 * reorthogonalisation on the device.  Non-Hermitian operators (Arpack.eigs'
 * complex path, which SA -> SR, LA -> LR, SM, LM; BE is rejected as ARPACK
 * rejects it): Krylov-Schur restarted Arnoldi, complex dtypes only, complex
 * Ritz values ordered as operators.jl:3393-3401.
 *   dtype  QS_F64/QS_C128: FP64 vectors; QS_F32/QS_C64: FP32-stored Krylov
 *          basis with FP64 dots, norms and tridiagonal solve.
 *   tol    <= 0 selects machine epsilon (ARPACK default)
 *   v0     NULL or Ns elements of the vector dtype (host)
 *   vecs   NULL or Ns*k elements of the vector dtype (host, column-major)
 *   vals   k doubles, or 2k (interleaved) when the operator is not Hermitian
 * Returns QS_ENOCONV (with the converged count in *nconv) if fewer than k
 * Ritz pairs converged within maxiter restarts. */
int qs_eigsh(qs_op *op, int dtype, int k, int which, double tol, int maxiter, int ncv,
             const void *v0, double *vals, void *vecs, int *nconv, int *n_matvec);

/* _krylov_expmv_times(op, v, scales; tol, krylov_dim) -- operators.jl:1805-1893:
 * out[:, j] = exp(scales[j] * op) * v for n_times complex scales.
 * dtype QS_C128 (FP64) or QS_C64 (FP32-stored basis, FP64 reductions).
 * v and out are host ComplexF64/ComplexF32 arrays (out: Ns x n_times). */
int qs_expm_multiply(qs_op *op, int dtype, const void *v, int64_t n_times, const double *scales,
                     double tol, int krylov_dim, void *out, int *n_matvec);

/* evolve(H, v0, t0, times; imag_time, max_step, rtol, atol) for a
 * Hamiltonian with dynamic terms, H(t) = h0 + sum_j f_j(t) dyn[j] --
 * operators.jl:2092-2128 with the adaptive Dormand-Prince step
 * _rk45_interval (operators.jl:1964-2071), Schroedinger equation
 * dv/dt = -i H(t) v (imag_time: -H(t) v, each state renormalised per
 * column).  The drive values come from `drive(t, n_dyn, coef, user)`, which
 * writes f_j(t) as 2 * n_dyn interleaved (re, im) doubles; it is called on the
 * calling thread once per stage time.  Every operator acts on the basis of
 * h0.  dtype QS_C128 (the reference's ComplexF64 state) or QS_C64.  v0: host,
 * Ns x ncols column-major; out: host, Ns x ncols x n_times.  times must be
 * sorted and not precede t0.  *n_steps (may be NULL) = attempted steps,
 * *n_matvec (may be NULL) = operator products. */
typedef void (*qs_drive_fn)(double t, int n_dyn, double *coef, void *user);
int qs_evolve_dynamic(qs_op *h0, int n_dyn, qs_op *const *dyn, qs_drive_fn drive, void *user, int dtype,
                      const void *v0, int64_t ncols, double t0, int64_t n_times, const double *times, int imag_time,
                      double rtol, double atol, double max_step, void *out, int64_t *n_steps, int64_t *n_matvec);

/* lanczos_full(op, v0, m; full_ortho, eps) and the coefficients of
 * lanczos_iter -- tools.jl:2742-2850: the reference's three-term recurrence
 * on the device.  v0: host Ns elements of dtype; eps <= 0 selects the machine
 * epsilon of dtype.  Outputs (host): alphas[m], betas[m-1], the
 * eigen(SymTridiagonal) values E[used] (ascending) and vectors V (used x used,
 * column-major), and, when Q_T is not NULL, the used Lanczos vectors as the
 * rows of Q_T (m x Ns row-major).  *used = rows actually built (early
 * breakdown at beta <= eps). */
int qs_lanczos(qs_op *op, int dtype, const void *v0, int m, int full_ortho, double eps, double *alphas,
               double *betas, double *E, double *V, void *Q_T, int *used);

/* The Gram products of the Gram-Schmidt step, G = V^H W -- the dot(basis,
 * residual) coefficients of _krylov_factorization (operators.jl:1697-1719),
 * for a block W of p vectors at once.  V: Ns x m, W: Ns x p, column-major
 * arrays of dtype; G: m x p interleaved complex (column-major), FP64
 * accumulation.  Bit 0 of on_mfma: the block kernel on the matrix cores (the
 * one the FP32-stream eigsh uses for its Rayleigh-Ritz products), else the
 * VALU Gram kernel of the Arnoldi step, one column of W at a time; bit 1: V
 * and W are device arrays (else host arrays, copied in). */
int qs_block_gram(int dtype, const void *V, int64_t n, int m, const void *W, int p, int on_mfma, double *G);

/* ------------------------------------------------ blocks of one parent space --
 * BlockOps' psi_0 lives in the basis built without the block keywords
 * (tools.jl:2447-2456).  kind 0: that parent is the block basis' own particle
 * sector, indexed by combinadic rank in ascending state order (all 2^L
 * states, indexed by the state, without nup); kind 1: the parent has no
 * particle number while the block has (blocks over nup), indexed by the
 * state.  Device pointers. */
int qs_basis_parent_size(const qs_basis *b, int kind, int64_t *n);

/* projector' * psi -- tools.jl:2453 (projection_matrix, symmetry_basis.jl:234-255):
 * out (Ns) from parent (parent size). */
int qs_basis_project(const qs_basis *b, int dtype, int kind, const void *parent, void *out);

/* parent (+)= projector * x for ncols columns -- _accumulate_projection!,
 * tools.jl:2488-2495 (x: Ns x ncols, parent: parent size x ncols, both
 * column-major).  Symmetry sectors need the parent lookup table (L <= 32). */
int qs_basis_expand(const qs_basis *b, int dtype, int kind, const void *x, void *parent, int64_t ncols,
                    int accumulate);

/* ------------------------------------------------------------ batches --------
 * Many operators (sectors) advanced in lockstep: one launch per Krylov step
 * for all of them.  Every job's result is bitwise the same whatever else is
 * in the batch.  Operators must take the two-bit hop kernel (see
 * qs_op_set_variant).  Vectors are device pointers; calls are synchronous. */
typedef struct qs_batch qs_batch;

/* No reference counterpart (the reference loops over blocks,
 * tools.jl:1864-1882); the operators must outlive the batch. */
int qs_batch_create(qs_op *const *ops, int n_ops, int dtype, qs_batch **out);
int qs_batch_destroy(qs_batch *b);

/* y_j = op[op_index[j]] x_j for n_jobs jobs in one launch per kernel class. */
int qs_batch_matvec(qs_batch *b, int n_jobs, const int32_t *op_index, const void *const *x, void *const *y);

/* _krylov_expmv_times (operators.jl:1805-1893) of n_jobs (operator, v,
 * time grid) jobs, as _block_batch_states runs them per block
 * (tools.jl:2500-2520): job j's grid is n_times[j] complex scales taken in
 * order from `scales` (interleaved re, im); out[j] is Ns_j x n_times[j]. */
int qs_batch_expm_multiply(qs_batch *b, int n_jobs, const int32_t *op_index, const void *const *v,
                           const int64_t *n_times, const double *scales, double tol, int krylov_dim,
                           void *const *out, int32_t *n_matvec);

/* evolve(::BlockOps, psi_0, t0, times) / block_expm with host vectors --
 * tools.jl:2445-2646: psi (host, parent size of kind[j] for every block)
 * is projected into each block (blocks whose projection has norm <= 1000 eps
 * are skipped, as in _active_block_entries), every active block runs
 * _krylov_expmv_times in one lockstep batch, and out (host, parent size x
 * n_times, column-major) receives sum_j P_j states_j.  dtype QS_C128 or QS_C64. */
int qs_blocks_evolve(qs_op *const *ops, const int32_t *kind, int n_ops, int dtype, const void *psi,
                     int64_t n_times, const double *scales, double tol, int krylov_dim, void *out,
                     int32_t *n_matvec);

/* ------------------------------------------------------------- serving -----
 * A request loop on one GPU: n_workers workers, each with its own HIP stream,
 * take eigsh / expm jobs first-in first-out; a job starts when a worker is
 * free and its Krylov working set fits the HBM budget beside the running
 * jobs', so requests that arrive while others run are admitted as slots and
 * memory free up and their kernels overlap on the device.  A job's result is
 * bitwise the same alone and in any mix.  No reference counterpart: the
 * reference answers one call at a time (eigsh operators.jl:3367-3408,
 * _krylov_expmv_times :1805-1893); every job runs exactly those. */
#define QS_JOB_EIGSH 0
#define QS_JOB_EXPM 1
typedef struct qs_job {
    int kind;               /* QS_JOB_EIGSH or QS_JOB_EXPM */
    qs_op *op;              /* must outlive the job */
    int dtype;
    int k, which, maxiter, ncv;   /* eigsh arguments as in qs_eigsh */
    double tol;             /* eigsh tol, or the expm tolerance */
    const void *v0;         /* host, Ns elements: eigsh start (may be NULL) or expm v (required); copied */
    int64_t n_times;        /* expm: number of complex scales */
    const double *scales;   /* expm: interleaved (re, im); copied */
    int krylov_dim;         /* expm */
    int want_vectors;       /* eigsh: keep the Ritz vectors (Ns x k) */
} qs_job;
typedef struct qs_server qs_server;

int qs_server_create(int n_workers, int64_t hbm_budget_bytes, qs_server **out);
/* Runs every queued job, then stops the workers. */
int qs_server_destroy(qs_server *s);
int qs_server_submit(qs_server *s, const qs_job *job, int64_t *ticket);
/* Blocks until the job is done and hands back its results once: eigsh vals
 * (k, or 2k interleaved when not Hermitian) and vectors in `out` (Ns x k) if
 * asked for; expm states in `out` (Ns x n_times).  seconds[3] = time queued,
 * time running, total.  Returns the job's own status. */
int qs_server_wait(qs_server *s, int64_t ticket, double *vals, void *out, int *nconv, int *n_matvec,
                   double *seconds);

#ifdef __cplusplus
}
#endif
#endif /* QSGPU_H */
