"""GPU tests of eigsh on non-Hermitian operators (qs_eigsh's Arnoldi path).

The reference passes a non-Hermitian QuantumLinearOperator to Arpack.eigs
(operators.jl:3383-3390): complex Ritz values for which = SR (asked as SA),
LR (LA), SM and LM, ordered by operators.jl:3393-3401, and the dense
decomposition for k >= Ns - 1 (:3379-3382).  Checked against numpy's dense
eigenvalues of the oracle's sector matrix P'HP: the k wanted values within
1e-9 of the dense ones and every Ritz pair's residual |A x - theta x| within
1e-9 |A| with |x| = 1.
"""
import numpy as np
import pytest

import oracle
import qsgpu

pytestmark = pytest.mark.gpu


def _terms(L, seed):
    rng = np.random.default_rng(seed)
    # asymmetric hopping (Hatano-Nelson), Ising coupling and complex fields
    t = [("+-", [(1.0, i, i % L + 1) for i in range(1, L + 1)]),
         ("-+", [(0.55, i, i % L + 1) for i in range(1, L + 1)]),
         ("zz", [(0.8, i, i % L + 1) for i in range(1, L + 1)])]
    t.append(("z", [(complex(rng.uniform(-0.5, 0.5), rng.uniform(-0.3, 0.3)), i) for i in range(1, L + 1)]))
    return t


def _key(which):
    return {"SA": lambda z: z.real, "LA": lambda z: -z.real, "SM": np.abs, "LM": lambda z: -np.abs(z)}[which]


CASES = [(12, 6, {}), (12, 5, {}), (14, 7, dict(kblock=1))]


def _uniform_terms(L):
    # translation invariant: a momentum sector keeps the operator block diagonal
    return [("+-", [(1.0, i, i % L + 1) for i in range(1, L + 1)]),
            ("-+", [(0.55, i, i % L + 1) for i in range(1, L + 1)]),
            ("zz", [(0.8 + 0.25j, i, i % L + 1) for i in range(1, L + 1)])]


# SM without a shift asks for interior Ritz values, which Arnoldi finds only
# slowly (as ARPACK's regular mode does): small sectors with a wide basis
SM_CASES = [(10, 3, {}), (10, 4, {}), (12, 6, dict(kblock=1))]


@pytest.mark.parametrize("which,case", [(w, c) for w in ("SA", "LA", "LM") for c in CASES] +
                         [("SM", c) for c in SM_CASES],
                         ids=["%s_L%d_n%d%s" % (w, c[0], c[1], "_k1" if c[2] else "")
                              for w in ("SA", "LA", "LM") for c in CASES] +
                         ["SM_L%d_n%d%s" % (c[0], c[1], "_k1" if c[2] else "") for c in SM_CASES])
def test_nonhermitian_eigsh_against_dense(gpu, case, which):
    L, nup, blocks = case
    print("non-Hermitian eigsh", case, which, flush=True)
    terms = _uniform_terms(L) if blocks else _terms(L, 5 + L)
    sec = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    A = oracle.Operator(sec, terms, check_symm=False).dense()
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms, check_herm=False, check_symm=False)
    assert not op.ishermitian()
    k = 4
    vals, vecs = qsgpu.eigsh(op, k=k, which=which, tol=1e-12, maxiter=2000, ncv=60 if which == "SM" else 40)
    assert vals.dtype == np.complex128 and vecs.shape == (b.Ns, k)
    dense = np.linalg.eigvals(A)
    key = _key(which)
    want = dense[np.argsort(key(dense), kind="stable")][:k]
    # the reported order is the requested one
    assert np.all(np.diff(key(vals)) >= -1e-9)
    for v in vals:
        assert np.min(np.abs(want - v)) <= 1e-9 * max(1.0, abs(v))
    nA = np.linalg.norm(A, 2)
    for i in range(k):
        x = vecs[:, i]
        assert abs(np.linalg.norm(x) - 1) <= 1e-10
        assert np.linalg.norm(A @ x - vals[i] * x) <= 1e-9 * nA


def test_nonhermitian_eigsh_fp32_stream_dense_fallback_and_errors(gpu):
    print("non-Hermitian eigsh fp32 / dense / errors", flush=True)
    L, nup = 12, 6
    terms = _terms(L, 17)
    sec = oracle.Sector(L, nup=nup, pauli=False)
    A = oracle.Operator(sec, terms, check_symm=False).dense()
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False)
    op = qsgpu.QuantumLinearOperator(b, terms, check_herm=False, check_symm=False)
    dense = np.linalg.eigvals(A)
    want = dense[np.argsort(dense.real, kind="stable")][:3]
    # ComplexF32-stored basis, FP64 reductions: FP32 accuracy
    v32 = qsgpu.eigsh(op, k=3, which="SA", tol=1e-6, maxiter=2000, ncv=40, dtype=np.complex64,
                      return_eigenvectors=False)
    assert np.abs(v32 - want).max() <= 2e-4 * np.abs(want).max()
    # a real dtype request is promoted to the complex path, as Arpack.eigs returns complex values
    v64 = qsgpu.eigsh(op, k=3, which="SA", tol=1e-12, maxiter=2000, ncv=40, dtype=np.float64,
                      return_eigenvectors=False)
    assert v64.dtype == np.complex128 and np.abs(v64 - want).max() <= 1e-9 * np.abs(want).max()
    with pytest.raises(qsgpu.ArgumentError, match="BE"):
        qsgpu.eigsh(op, k=2, which="BE")
    # k >= Ns - 1: the dense decomposition (operators.jl:3379-3382)
    Ls, ns = 6, 1
    ts = _terms(Ls, 3)
    As = oracle.Operator(oracle.Sector(Ls, nup=ns, pauli=False), ts, check_symm=False).dense()
    bs = qsgpu.SpinBasis1D(Ls, nup=ns, pauli=False)
    ops = qsgpu.QuantumLinearOperator(bs, ts, check_herm=False, check_symm=False)
    vals, vecs = qsgpu.eigsh(ops, k=bs.Ns - 1, which="LM")
    d = np.linalg.eigvals(As)
    d = d[np.argsort(-np.abs(d), kind="stable")][:bs.Ns - 1]
    assert np.abs(vals - d).max() <= 1e-12
    for i in range(bs.Ns - 1):
        assert np.linalg.norm(As @ vecs[:, i] - vals[i] * vecs[:, i]) <= 1e-12
