#!/usr/bin/env python3
"""Benchmark: sector H·v throughput on one MI355X.

Workload (default): J1-J2 XXZ chain, L=28, nup=14, k=0, p=+1, z=+1
(N = 361,270), J1=1, J2=0.4, Delta=1, FP64 matrix-free matvec.  A step is one
matvec over the whole sector with x and y resident in HBM.  `value` is
algorithmic bytes per matvec / time, with
    bytes = N * (8 read x + 8 write y + 4 stored state + 8 stored 1/sqrt(D)).
Beside it: the FP64 Lanczos ground state (k=1, tol 1e-10) of the same
sector, the L=32 z+ FP32-stream / FP64 matvec pair (N = 300,540,195,
implicit basis; bytes = N * 2 * element), the batched block evolution
(L=24, all momentum blocks, Neel quench), the serving stream (Poisson, mixed eigsh /
evolve, solves/s and p50/p99 latency) and the CPU baselines: the oracle's restatement of QuSpin.jl's
whole QuantumLinearOperator product (lift, parent-space apply, project; operators.jl:2591-2623) on
the host's threads (median of warm products), the same apply loop on one thread over a bounded
sample, and C1's CSR SpMV and ARPACK eigsh beside the GPU's.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-c3] [--no-c4] [--no-c5]
                       [--no-eigsh]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "quspin.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def xxz_terms(L, J=1.0, delta=1.0, J2=0.0):
    b1 = [(i, i % L + 1) for i in range(1, L + 1)]
    terms = [("+-", [(J / 2, i, j) for i, j in b1]), ("-+", [(J / 2, i, j) for i, j in b1]),
             ("zz", [(J * delta, i, j) for i, j in b1])]
    if J2:
        b2 = [(i, (i + 1) % L + 1) for i in range(1, L + 1)]
        terms += [("+-", [(J2 / 2, i, j) for i, j in b2]), ("-+", [(J2 / 2, i, j) for i, j in b2]),
                  ("zz", [(J2 * delta, i, j) for i, j in b2])]
    return terms


def profiled_traffic(key):
    """HBM bytes per launch from the committed rocprofv3 PMC summary (FETCH_SIZE
    and WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f)[key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def variant_name(v):
    import qsgpu
    if v == qsgpu.VARIANT_SELL:
        return "k_sell_spmv (assembled SELL-64 rows)"
    if v >= qsgpu.VARIANT_SPLIT + 4:
        return "k_matvec_split2 x 2 passes (lean split kernel, tiled two-pass form, tile budget %d MB)" % (
            1 << (v - qsgpu.VARIANT_SPLIT - 4))
    if v == qsgpu.VARIANT_SPLIT + 3:
        return "k_matvec_split2 (lean split kernel: table-driven low / cross hops, scalar high hops)"
    if v >= qsgpu.VARIANT_SPLIT:
        fewer = v - qsgpu.VARIANT_SPLIT
        return "k_matvec_split (high-word blocks, low word in LDS%s)" % (
            "; %d fewer resident workgroups" % fewer if fewer else "")
    if v == qsgpu.VARIANT_HOP + 3:
        return "k_matvec_hop (1 wave per row tile, pre-scaled x)"
    if v >= qsgpu.VARIANT_HOP:
        return "k_matvec_hop (%d wave(s) per row tile)" % (1 << (v - qsgpu.VARIANT_HOP))
    return "general kernel, %d lane(s) per row" % v


def time_matvec(op, x, y, steps, warmup):
    """Average matvec time (s) with HIP events on the stream the kernel runs on."""
    import torch
    stream = torch.cuda.current_stream()
    op.set_stream(stream.cuda_stream)
    for _ in range(warmup):
        op.matvec_device(x, y)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        op.matvec_device(x, y)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(end) * 1e-3 / steps, wall


def bench_c4(n_times):
    """Batched form: quench from the Neel state in every momentum block (p-blocks at k = 0, pi) of the
    L=24 nup=12 XXZ chain (J=1, Delta=2), exp(-iHt) on a grid of t in [0, 20], all blocks in one
    lockstep engine; ComplexF64 and ComplexF32-stream forms."""
    import numpy as np
    import torch
    import qsgpu
    from qsgpu import serve
    L, nup = 24, 12
    print("bench: C4 blocks", file=sys.stderr, flush=True)
    blocks = []
    for k in range(L):
        blocks += [dict(kblock=k, pblock=1), dict(kblock=k, pblock=-1)] if k in (0, L // 2) else [dict(kblock=k)]
    tb = time.perf_counter()
    ops = qsgpu.BlockOps(blocks, serve.chain_terms(L, 0.0, 2.0), [], qsgpu.SpinBasis1D, (L,), np.complex128,
                         basis_kwargs=dict(nup=nup, pauli=False), compute_all_blocks=True)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - tb
    neel = sum(1 << i for i in range(0, L, 2))
    rank, kk, p, st = 0, 1, 0, neel
    while st:
        if st & 1:
            rank += math.comb(p, kk)
            kk += 1
        st >>= 1
        p += 1
    psi = np.zeros(math.comb(L, nup), dtype=np.complex128)
    psi[rank] = 1.0
    times = np.linspace(0.0, 20.0, n_times)
    out = {"blocks": len(ops.basis_dict), "parent_dim": int(psi.size), "n_times": int(n_times), "t_max": 20.0,
           "build_s": round(t_build, 3)}
    ref = None
    for dt, name, tol in ((np.complex128, "c128", 1e-12), (np.complex64, "c64", 1e-6)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = ops.evolve(psi, 0.0, times, dtype=dt, tol=tol, device=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        mv = int(sum(ops.last_n_matvec))
        rows = sum(int(ops.basis_dict[k].Ns) for k in ops.basis_dict)
        # the states stay in HBM; the copy of all n_times parent vectors to (pageable) host
        # memory, which evolve() without device=True adds, is timed apart
        t1 = time.perf_counter()
        host = st.cpu()
        el_copy = time.perf_counter() - t1
        out[name] = {"seconds": round(el, 3), "block_matvecs": mv, "tol": tol,
                     "host_copy_s": round(el_copy, 3), "host_copy_GB": round(host.numel() * host.element_size() / 1e9, 2)}
        if ref is None:
            ref = st
        else:
            out[name]["max_abs_err_vs_c128"] = float((st.to(torch.complex128) - ref).abs().max())
        del host
    out["rows_all_blocks"] = rows
    return out


def bench_c3(args, skip_eigsh=False):
    """Precision form's matvec pair: L=32 nup=16 z+ (N = 300,540,195), FP32 stream and FP64, each
    with the autotuned (or pinned) kernel variant; then, unless skipped, the k=4 eigensolve."""
    import numpy as np
    import torch
    import qsgpu
    print("bench: C3 L=32 z+ matvec pair", file=sys.stderr, flush=True)
    b3 = qsgpu.SpinBasis1D(32, nup=16, pauli=False, zblock=1)
    # disordered fields h_i ~ U[-2, 2]: odd under spin inversion, so the
    # sector's group average P'HP cancels them exactly (check_symm=false)
    h = np.random.default_rng(0xC3).uniform(-2.0, 2.0, 32)
    terms3 = xxz_terms(32, J=1.0, delta=1.0) + [("z", [(float(h[i]), i + 1) for i in range(32)])]
    op3 = qsgpu.QuantumLinearOperator(b3, terms3, check_symm=False)
    n3 = b3.Ns
    out = {"sector_dim": n3}
    for dt, name, esz in ((torch.float32, "fp32", 4), (torch.float64, "fp64", 8)):
        x3 = torch.rand(n3, dtype=dt, device="cuda")
        y3 = torch.zeros_like(x3)
        if args.c3_variant:
            op3.set_variant(args.c3_variant)
            lpr3 = args.c3_variant
        else:
            lpr3, _ = op3.tune(np.float32 if esz == 4 else np.float64, repeats=2)
        t3, _ = time_matvec(op3, x3, y3, max(3, args.steps // 10), 2)
        b3_alg = n3 * 2 * esz   # read x once, write y once; the z+ basis is implicit (bit 31 clear)
        out[name] = {"ms": round(t3 * 1e3, 4), "GBps": round(b3_alg / t3 / 1e9, 2), "variant": lpr3,
                     "kernel": variant_name(lpr3),
                     "frac": round(b3_alg / t3 / 1e9 / HBM_PEAK_GBPS, 4), "bytes_per_matvec": b3_alg,
                     "traffic": profiled_traffic("c3_" + name)}
        del x3, y3
        torch.cuda.empty_cache()
    out["fp32_over_fp64"] = round(out["fp64"]["ms"] / out["fp32"]["ms"], 3)
    eig = None if skip_eigsh else bench_c3_eigsh(op3, n3)
    del op3, b3
    torch.cuda.empty_cache()
    return out, eig


def bench_gram(n=1 << 23, m=32, p=32, repeats=5):
    """The block Gram product G = V^H W of the Krylov path on the matrix cores (k_gram_mfma,
    v_mfma_f64_16x16x4f64, FP32 widened to FP64), V: n x m, W: n x p resident on the device; wall
    time per call (launches, the cross-block reduction and the m x p copy to the host included)."""
    import ctypes
    import numpy as np
    import torch
    from qsgpu import _lib
    print("bench: block Gram on MFMA", file=sys.stderr, flush=True)
    out = {"n": n, "m": m, "p": p}
    for dt, name, code in ((torch.float64, "fp64", _lib.QS_F64), (torch.float32, "fp32", _lib.QS_F32)):
        g = torch.Generator(device="cpu").manual_seed(3)
        V = torch.randn(m, n, generator=g, dtype=torch.float64).to(dt).cuda()   # column j = row j here
        W = torch.randn(p, n, generator=g, dtype=torch.float64).to(dt).cuda()
        G = np.zeros(2 * m * p)
        call = lambda: _lib.lib().qs_block_gram(code, ctypes.c_void_p(V.data_ptr()), n, m,
                                                 ctypes.c_void_p(W.data_ptr()), p, 3, G.ctypes.data)
        assert call() == 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            call()
        t = (time.perf_counter() - t0) / repeats
        exact = (V.double() @ W.double().T).cpu().numpy()
        got = G[0::2].reshape((m, p), order="F")
        esz = 8 if dt == torch.float64 else 4
        out[name] = {"ms": round(t * 1e3, 4), "TFLOPs": round(2.0 * n * m * p / t / 1e12, 3),
                     "GBps": round(n * (m + p) * esz / t / 1e9, 1),
                     "max_rel_err_vs_fp64": float(np.abs(got - exact).max() / np.abs(exact).max())}
        del V, W
        torch.cuda.empty_cache()
    out["note"] = "tall-skinny: 2 m p / ((m + p) s) flop per byte, so HBM-bound; FP64 MFMA peak 78.6 TFLOP/s"
    return out


def bench_c3_eigsh(op3, n3):
    """Precision form: the k=4 lowest eigenvalues of the L=32 z+ sector in the FP32-stream form
    (FP32 Krylov basis, FP64 dots / norms / tridiagonal solve, FP64 Rayleigh quotients) and in
    FP64; each eigenvalue's error against the FP64 GPU run, and each FP64 Ritz vector's residual."""
    import numpy as np
    import torch
    import qsgpu
    print("bench: C3 eigsh k=4 FP64 / FP32 stream", file=sys.stderr, flush=True)
    v0 = np.random.default_rng(0x5EED).standard_normal(n3)
    out = {"k": 4, "which": "SA", "tol": 1e-10}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e64, vec = qsgpu.eigsh(op3, k=4, which="SA", tol=1e-10, v0=v0)
    out["seconds_fp64"] = round(time.perf_counter() - t0, 3)
    out["matvecs_fp64"] = qsgpu.eigsh.last_n_matvec
    y = torch.empty(n3, dtype=torch.float64, device="cuda")
    res = []
    for i in range(4):
        v = torch.from_numpy(np.ascontiguousarray(vec[:, i])).cuda()
        op3.matvec_device(v, y)
        res.append((torch.linalg.vector_norm(y - float(e64[i]) * v) / torch.linalg.vector_norm(v)).item())
        del v
    del vec, y
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    e32 = qsgpu.eigsh(op3, k=4, which="SA", tol=1e-10, v0=v0, dtype=np.float32, return_eigenvectors=False)
    out["seconds_fp32"] = round(time.perf_counter() - t0, 3)
    out["matvecs_fp32"] = qsgpu.eigsh.last_n_matvec
    out["E_fp64"] = [float(e) for e in e64]
    out["E_fp32_stream"] = [float(e) for e in e32]
    out["abs_err_fp32_vs_fp64"] = [abs(float(a) - float(b)) for a, b in zip(e32, e64)]
    out["residual_fp64"] = res
    return out


def _median(v):
    v = sorted(v)
    n = len(v)
    return v[n // 2] if n % 2 else 0.5 * (v[n // 2 - 1] + v[n // 2])


def bench_cpu(L, terms, bytes_alg, t_mv, args):
    """CPU baselines: the oracle's restatement of QuSpin.jl's CPU algorithms on this host's cores.
    (1) the whole C2 product of the reference's QuantumLinearOperator (lift x with P, parent-space
    apply with one hash lookup per hit, project with P', operators.jl:2591-2623) on the host's
    threads (at most 16, one GPU's share), median of warm products; (2) the same apply loop on one
    thread over a bounded sample of parent states; (3) C1: the CSR SpMV of sparse_formats.jl:110-146
    and eigsh(k=4, :SA) through ARPACK (scipy.sparse.linalg.eigsh, the ARPACK-ng family Arpack.jl
    wraps), beside the GPU's products and eigsh on the same C1 operator."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    import oracle
    import qsgpu
    out = {}
    nth = max(1, min(os.cpu_count() or 1, 16))
    print("bench: CPU baseline, whole C2 product on %d threads" % nth, file=sys.stderr, flush=True)
    reps = max(1, args.cpu_repeats)
    ns, tt = oracle.bench_qlo_matvec(L, 14, False, terms, nth, reps, kblock=0, pblock=1, zblock=1)
    warm = tt
    tot = [float(a) for a in warm.sum(axis=1)]
    med = _median(tot)
    out["cpu_baseline"] = {
        "value": round(bytes_alg / med / 1e9, 9), "unit": "GB/s", "cores": nth, "kind": "port",
        "sample": "the whole C2 product (N=%d; lift P x, parent-space apply over all %d parent states with a "
                  "hash lookup per hit, project P'), median of %d products (work vectors touched beforehand): %.2f s "
                  "(lift %.3f s, apply %.2f s, project %.3f s); no extrapolation"
                  % (ns, math.comb(L, 14), reps, med, _median(list(warm[:, 0])), _median(list(warm[:, 1])),
                     _median(list(warm[:, 2]))),
        "seconds_per_product": round(med, 4), "samples": reps,
        "speedup_gpu_over_cpu": round(med / t_mv, 1),
    }
    print("bench: CPU baseline, 1 thread, %d-state sample" % args.cpu_sample, file=sys.stderr, flush=True)
    npar = math.comb(L, 14)
    ts = oracle.bench_parent_apply(L, 14, False, terms, args.cpu_sample, 5)
    t_full = ts * npar / args.cpu_sample
    out["cpu_baseline_1thread"] = {
        "value": round(bytes_alg / t_full / 1e9, 9), "unit": "GB/s", "cores": 1, "kind": "port",
        "sample": "the parent-space apply loop (operators.jl:2554-2589, hash lookup) over the first %d of %d "
                  "parent states, median of 5: %.2f s, scaled by %d / %d to one product (%.1f s; lift and "
                  "project, about %.2f s on %d threads above, not included)"
                  % (args.cpu_sample, npar, ts, npar, args.cpu_sample, t_full,
                     _median(list(warm[:, 0] + warm[:, 2])), nth),
        "seconds_per_product_extrapolated": round(t_full, 2), "samples": 5,
        "speedup_gpu_over_cpu": round(t_full / t_mv, 1),
    }
    # C1: XXZ L=16 nup=8 pauli=false, J=1, Delta=0.5
    print("bench: C1 CSR SpMV and ARPACK eigsh (CPU) against the GPU", file=sys.stderr, flush=True)
    L1 = 16
    b1 = [(i, i % L1 + 1) for i in range(1, L1 + 1)]
    terms1 = [("+-", [(0.5, i, j) for i, j in b1]), ("-+", [(0.5, i, j) for i, j in b1]),
              ("zz", [(0.5, i, j) for i, j in b1])]
    sec1 = oracle.Sector(L1, nup=8, pauli=False)
    csr = oracle.CSR(sec1, terms1)
    rowptr, col, val = csr.export()
    A = sp.csr_matrix((val, col, rowptr), shape=(sec1.Ns, sec1.Ns))
    rng = np.random.default_rng(0xC1)
    x1 = rng.standard_normal(sec1.Ns)
    x1 /= np.linalg.norm(x1)
    y1 = np.zeros(sec1.Ns)
    csr.matvec(x1, y=y1)
    spmv = []
    for _ in range(20):
        t0 = time.perf_counter()
        csr.matvec(x1, y=y1)
        spmv.append(time.perf_counter() - t0)
    eig_cpu = []
    for _ in range(5):
        t0 = time.perf_counter()
        w = spla.eigsh(A, k=4, which="SA", v0=x1, return_eigenvectors=False)
        eig_cpu.append(time.perf_counter() - t0)
    bg = qsgpu.SpinBasis1D(L1, nup=8, pauli=False)
    opg = qsgpu.QuantumLinearOperator(bg, terms1)
    xd = torch.from_numpy(x1).cuda()
    yd = torch.zeros_like(xd)
    t_gpu_mv, _ = time_matvec(opg, xd, yd, 200, 10)
    eig_gpu = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        wg = qsgpu.eigsh(opg, k=4, which="SA", v0=x1, return_eigenvectors=False)
        eig_gpu.append(time.perf_counter() - t0)
    c1_bytes = sec1.Ns * 16
    out["c1"] = {
        "sector_dim": int(sec1.Ns), "csr_nnz": int(csr.nnz),
        "cpu_csr_spmv_s": _median(spmv), "cpu_csr_spmv_samples": 20, "cpu_threads": 1,
        "gpu_matvec_s": round(t_gpu_mv, 8), "gpu_matvec_GBps": round(c1_bytes / t_gpu_mv / 1e9, 3),
        "cpu_arpack_eigsh_k4_s": round(_median(eig_cpu), 5), "cpu_eigsh_samples": 5,
        "gpu_eigsh_k4_s": round(_median(eig_gpu), 5), "gpu_eigsh_samples": 5,
        "E_cpu_arpack": sorted(float(v) for v in w), "E_gpu": [float(v) for v in wg],
        "max_abs_dE": float(np.max(np.abs(np.sort(w) - np.asarray(wg)))),
        "note": "CPU: the oracle's CSR SpMV (sparse_formats.jl:110-146 restated) and scipy's ARPACK; GPU: the "
                "matrix-free product and the device thick-restart Lanczos, both on this C1 operator",
    }
    return out


def bench_c5(n_requests, n_workers, assemble=True):
    """Serving form: Poisson stream (mean 200 req/s, seeded) of 60% eigsh k=1..4 / 40% evolve to
    t ~ U[1,10] on XXZ / J1-J2 chains, L in {16,20,24,28}, nup=L/2, random sectors, HBM budget 200 GB.
    Sector bases are built (cached) before the clock starts; each request's operator is built at
    arrival.  Latency = completion - scheduled arrival."""
    import numpy as np
    import qsgpu
    from qsgpu import serve
    print("bench: C5 serving", file=sys.stderr, flush=True)
    reqs = serve.poisson_workload(n_requests, rate=200.0, seed=2026)
    warm = serve.RequestBuilder()
    for r in reqs:
        try:
            warm.operator(r)
        except qsgpu.ArgumentError:
            r["skip"] = True
    # operators are built (and, for sectors up to 4M rows, assembled) at arrival, on the clock
    builder = serve.RequestBuilder(assemble=assemble)
    builder.bases = warm.bases
    done = []
    with serve.Server(n_workers=n_workers, hbm_budget=200e9) as srv:
        start = time.perf_counter()
        tickets = []
        for r in reqs:
            if r.get("skip"):
                continue
            d = start + r["arrival"] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            op = builder.operator(r)
            if r["kind"] == "eigsh":
                t = srv.submit_eigsh(op, k=min(r["k"], op.Ns), which="SA", tol=1e-10, return_eigenvectors=True)
            else:
                t = srv.submit_evolve(op, builder.initial_state(r, op), 0.0, [r["t"]])
            tickets.append((r, t, op))
        for r, t, op in tickets:
            res = srv.wait(t)
            done.append((r, res, time.perf_counter() - start, op))
    lat = np.array([c - r["arrival"] for r, _, c, _ in done])
    span = max(c for _, _, c, _ in done)
    # eigenvalue error of every served eigsh request, off the clock: for a
    # Hermitian H and a unit Ritz vector v with value E, some eigenvalue lies
    # within ||H v - E v|| of E
    err = []
    for r, res, _, op in done:
        if r["kind"] != "eigsh":
            continue
        for i, e in enumerate(res["vals"]):
            v = res["vecs"][:, i]
            v = v / np.linalg.norm(v)
            err.append(float(np.linalg.norm(op.matvec(v) - e * v)))
    err = np.array(err)
    return {"requests": len(done), "rate_per_s": 200.0, "workers": n_workers, "hbm_budget_bytes": 200e9,
            "solves_per_s": round(len(done) / span, 2), "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4),
            "mean_run_s": round(float(np.mean([x["seconds"][1] for _, x, _, _ in done])), 4),
            "eigsh_tol": 1e-10, "evolve_tol": 1e-12, "assembled_operators": bool(assemble),
            "eigenvalue_error_bound_max": float(err.max()) if err.size else None,
            "eigenvalue_error_bound_p99": float(np.percentile(err, 99)) if err.size else None,
            "eigenvalues_checked": int(err.size),
            "error_note": "per served eigenvalue, ||H v - E v|| of its unit Ritz vector (some eigenvalue lies "
                          "that close, H Hermitian); evolve requests run to the Krylov error estimate evolve_tol"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c3-eigsh", action="store_true")
    ap.add_argument("--no-eigsh", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--c3-only", action="store_true",
                    help="only the L=32 z+ matvec pair (profiling runs); implies --no-c3-eigsh unless --c3-eigsh")
    ap.add_argument("--c3-eigsh", action="store_true", help="with --c3-only: also the k=4 eigensolve")
    ap.add_argument("--c3-variant", type=int, default=0, help="pin the C3 kernel variant instead of autotuning")
    ap.add_argument("--gram-only", action="store_true", help="only the block Gram product on the matrix cores")
    ap.add_argument("--c4-times", type=int, default=200)
    ap.add_argument("--c5-requests", type=int, default=200)
    ap.add_argument("--c5-workers", type=int, default=8)
    ap.add_argument("--c5-matrix-free", action="store_true", help="serve without assembling the operators")
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--cpu-repeats", type=int, default=3, help="warm products of the whole-sector CPU baseline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the library runs on exactly one GPU; sharding across GPUs is not implemented "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import qsgpu

    dev = qsgpu.device_info()
    if args.gram_only:
        print(json.dumps({"gram_mfma": bench_gram()}))
        return
    if args.c3_only:
        res = bench_c3(args, skip_eigsh=not args.c3_eigsh)
        print(json.dumps({"c3_matvec": res[0], "c3_eigsh": res[1]}))
        return
    L = 28
    print("bench: building C2 sector L=28 nup=14 k=0 p=+1 z=+1", file=sys.stderr, flush=True)
    tb = time.perf_counter()
    basis = qsgpu.SpinBasis1D(L, nup=14, pauli=False, kblock=0, pblock=1, zblock=1)
    torch.cuda.synchronize()
    t_basis = time.perf_counter() - tb
    terms = xxz_terms(L, J=1.0, delta=1.0, J2=0.4)
    op = qsgpu.QuantumLinearOperator(basis, terms)
    N = basis.Ns
    g = torch.Generator(device="cpu").manual_seed(1234)
    x = torch.randn(N, generator=g, dtype=torch.float64).cuda()
    x /= x.norm()
    y = torch.zeros_like(x)
    bytes_alg = N * (8 + 8 + 4 + 8)
    lpr, _ = op.tune(np.float64, repeats=5)
    print("bench: timing %d matvecs (N=%d, lanes/row %d)" % (args.steps, N, lpr), file=sys.stderr, flush=True)
    t_mv, wall = time_matvec(op, x, y, args.steps, args.warmup)
    gbps = bytes_alg / t_mv / 1e9
    result = {
        "metric": "sector H·v GB/s (% HBM peak)",
        "value": round(gbps, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mv * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian vector; Hamiltonian from closed-form couplings)",
        "config": {"workload": "C2: J1-J2 XXZ L=28 nup=14 k=0 p=+1 z=+1, N=%d, FP64 matrix-free matvec" % N,
                   "L": L, "nup": 14, "J1": 1.0, "J2": 0.4, "Delta": 1.0, "sector_dim": N,
                   "bytes_per_matvec": bytes_alg, "basis_build_s": round(t_basis, 4), "device": dev["name"],
                   "lanes_per_row": lpr},
        "roofline": {"bound": "hbm", "achieved": round(gbps, 3), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / HBM_PEAK_GBPS, 5), "traffic": profiled_traffic("c2_fp64"),
                     "kernel": "k_prescale + k_matvec_hop<double, real, parent table, pre-scaled x> "
                               "(variant %d)" % lpr,
                     "note": "one step = k_prescale (x/sqrt(D) to an FP64 buffer) + the hop kernel; per "
                             "neighbour: colex rank update, parent-table gather, one gather of the pre-scaled "
                             "x; achieved counts the whole step; the working set sits in L2/MALL at this N, "
                             "so the bound is gather issue and latency, not HBM"},
    }
    # the same sector matrix assembled once on the device (the reference's
    # static_fmt=:csr path): plain products read the stored rows
    ta = time.perf_counter()
    op_a = qsgpu.QuantumLinearOperator(basis, terms).assemble()
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - ta
    op_a.set_variant(qsgpu.VARIANT_SELL)
    t_sell, _ = time_matvec(op_a, x, y, args.steps, args.warmup)
    result["c2_assembled"] = {"ms": round(t_sell * 1e3, 5), "assemble_s": round(t_asm, 4),
                              "GBps_algorithmic": round(bytes_alg / t_sell / 1e9, 3),
                              "speedup_vs_matrix_free": round(t_mv / t_sell, 3),
                              "kernel": "k_sell_spmv (SELL-64 rows assembled by the pull kernel's enumeration)"}
    del op_a
    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        result["copy_calibration"] = {"bytes_read": src.numel() * 8, "bytes_written": dst.numel() * 8, "launches": 3}
        del src, dst
    if not args.no_eigsh:
        print("bench: FP64 Lanczos ground state", file=sys.stderr, flush=True)
        v0 = np.random.default_rng(0x51a7).standard_normal(N)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e64, _ = qsgpu.eigsh(op, k=1, which="SA", tol=1e-10, v0=v0, return_eigenvectors=True)
        t_e64 = time.perf_counter() - t0
        nmv64 = qsgpu.eigsh.last_n_matvec
        t0 = time.perf_counter()
        e32, _ = qsgpu.eigsh(op, k=1, which="SA", tol=1e-10, v0=v0, dtype=np.float32)
        t_e32 = time.perf_counter() - t0
        nmv32 = qsgpu.eigsh.last_n_matvec
        result["eigsh"] = {"E0_fp64": float(e64[0]), "seconds_fp64": round(t_e64, 4), "matvecs_fp64": nmv64,
                           "E0_fp32_stream": float(e32[0]), "seconds_fp32": round(t_e32, 4),
                           "matvecs_fp32": nmv32, "abs_err_fp32_vs_fp64": abs(float(e32[0]) - float(e64[0]))}
    if not args.no_c3:
        result["c3_matvec"], c3e = bench_c3(args, skip_eigsh=args.no_c3_eigsh)
        if c3e is not None:
            result["c3_eigsh"] = c3e
    result["gram_mfma"] = bench_gram()
    if not args.no_c4:
        result["c4_block_evolve"] = bench_c4(args.c4_times)
    if not args.no_c5:
        result["c5_serving"] = bench_c5(args.c5_requests, args.c5_workers, assemble=not args.c5_matrix_free)
    if not args.no_cpu:
        result.update(bench_cpu(L, terms, bytes_alg, t_mv, args))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
