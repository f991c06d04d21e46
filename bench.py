"""iSVD solves/s on one MI355X: A 100000 x 10000, k=20, p=12, N=64 (Y is 100000 x 2048), KN tol
1e-4, SVD former.  The plain FP64 path (A FP64) and the mixed path (the same A stored FP32, BF16
MFMA sketch with FP32 accumulation, FP64 Gram/eig/KN/SVD) run on the same Philox Omega; `value`
is the mixed path's solves/s, reported beside the FP64 path's and the sigma relative error
between them.  Prints one JSON line.

A is synthetic: A = U50 diag(0.9^i) V50' + E with Haar U50 (m x 50), V50 (n x 50) and
E_ij = 1e-3 g_ij / (sqrt(m) + sqrt(n)), g ~ N(0,1) (||E||_2 ~ 1e-3); it is resident in HBM
before timing starts.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK = {"bf16": 2.5e15, "f64": 78.6e12, "hbm": 8e12}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_factors(m, n, r, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return U, V


def timed_solves(mc, solver, A, steps, warmup):
    """Times `steps` solves (after `warmup`); also returns, per solve, the HIP-event durations of
    the sketch GEMM and Gram GEMM launches and of the four stages."""
    for _ in range(warmup):
        solver.solve(A)
    mc.synchronize()
    kern, stages = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.solve(A)
        times, _ = solver.stage_times()
        kern.append(solver.kernel_times())
        stages.append(times)
    mc.synchronize()
    dt = time.perf_counter() - t0
    return dt, np.array(kern), np.array(stages)


def load_traffic():
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a
    plain copy of known size), committed under profiles/; None when absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def cpu_baseline(m, n, k, p, N, iters, budget_rows):
    """The oracle (single-threaded C restatement) on a bounded row sample of the same workload:
    every stage is linear in m except l x l work, so solve time ~ sample time * m / m_s, with the
    integrator scaled to the GPU's iteration count."""
    import oracle
    l = k + p
    ms = budget_rows
    rng = np.random.default_rng(7)
    U, V = make_factors(ms, n, 50, 8)
    A = (U * 0.9 ** np.arange(50)) @ V.T + 1e-3 * rng.standard_normal((ms, n)) / (np.sqrt(m) + np.sqrt(n))
    om = rng.standard_normal((n, N * l))
    t = time.perf_counter()
    Y = oracle.sketch(A, om, a_colmajor=True)
    t_sk = time.perf_counter() - t
    t = time.perf_counter()
    Q, _ = oracle.gramian_orthogonalize(Y, l, N)
    t_or = time.perf_counter() - t
    t = time.perf_counter()
    Qb, it_s = oracle.kn_rowblock(Q, l, N, 8, 0.0)  # 8 fixed iterations
    t_it = (time.perf_counter() - t) / it_s
    t = time.perf_counter()
    oracle.svd_former(A, Qb, k, a_colmajor=True)
    t_fo = time.perf_counter() - t
    scale = m / ms
    total = (t_sk + t_or + t_fo) * scale + t_it * scale * iters
    sample_s = t_sk + t_or + t_it * it_s + t_fo
    return {"value": 1.0 / total, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "oracle pipeline (sketch, Gramian orthogonalizer, 8 KN iterations, SVD "
                      "former) on %d of %d rows (%.1f s of CPU), scaled linearly in m, KN to %d "
                      "iterations" % (ms, m, sample_s, iters),
            "seconds_per_solve": total}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=100000)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--p", type=int, default=12)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--cpu-rows", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed", default="bf16", choices=["bf16", "bf16x3"])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the iSVD path here is built for one MI355X "
              "(multi-GPU sharding is a design note, not built); rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import mcnla_amd as mc
    if mc.device_count() < 1:
        raise SystemExit("no GPU visible")
    m, n, k, p, N = args.m, args.n, args.k, args.p, args.N
    l = k + p
    log("bench: A %d x %d, k=%d p=%d N=%d" % (m, n, k, p, N))
    U, V = make_factors(m, n, 50, 1)
    sig = 0.9 ** np.arange(50)
    noise = 1e-3 / (np.sqrt(m) + np.sqrt(n))
    A64 = mc.synth_lowrank(m, n, U, V, sig, noise, seed=1, dtype=np.float64, order="F")
    A32 = mc.synth_lowrank(m, n, U, V, sig, noise, seed=1, dtype=np.float32, order="F")
    log("bench: A resident (FP64 %.1f GB, FP32 %.1f GB)" % (A64.nbytes / 1e9, A32.nbytes / 1e9))
    P = mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)
    prec = mc.PREC_MIXED_BF16 if args.mixed == "bf16" else mc.PREC_MIXED_BF16X3
    s64 = mc.Solver(P, seed=2)
    smx = mc.Solver(P, seed=2, precision=prec)

    fp64_steps = args.fp64_steps if args.fp64_steps is not None else max(3, args.steps // 2)
    dt64, k64, st64 = timed_solves(mc, s64, A64, fp64_steps, 1)
    g64 = k64[:, 0]
    log("bench: fp64 %.2f ms/solve (sketch gemm %.2f ms, gram gemm %.2f ms), iterations %d"
        % (dt64 / fp64_steps * 1e3, g64.mean() * 1e3, k64[:, 1].mean() * 1e3, s64.iteration()))
    dtmx, kmx, stmx = timed_solves(mc, smx, A32, args.steps, args.warmup)
    gmx, grmx = kmx[:, 0], kmx[:, 1]
    log("bench: mixed %.2f ms/solve (sketch gemm %.2f ms, gram gemm %.2f ms), iterations %d"
        % (dtmx / args.steps * 1e3, gmx.mean() * 1e3, grmx.mean() * 1e3, smx.iteration()))
    S64, Smx = s64.vectorS(), smx.vectorS()
    rel = float(np.max(np.abs(Smx - S64) / S64))
    Q64, Qmx = s64.matrixQbar(), smx.matrixQbar()
    # ||Q Q' - Qr Qr'||_2 = sin of the largest principal angle
    c = np.linalg.svd(Q64.T @ Qmx, compute_uv=False)
    dist = float(np.sqrt(max(0.0, 1.0 - c.min() ** 2)))

    nl = N * l
    flops_bf16 = 2.0 * m * n * nl * (1 if args.mixed == "bf16" else 3)
    flops_f64 = 2.0 * m * n * nl
    flops_gram = 1.0 * m * nl * (nl + 1)  # lower triangle with diagonal, 2 flops per product
    ach_mx = flops_bf16 / gmx.mean()
    ach_64 = flops_f64 / g64.mean()
    traffic = load_traffic()
    kernels = {
        "sketch_gemm": {"kernel": "sketch_blk_kernel (Y = A*Omega, BF16 MFMA, FP32 accumulate)"
                        if args.mixed == "bf16" else
                        "sketch_bf16_kernel<3> (Y = A*Omega, BF16x3 MFMA, FP32 accumulate)",
                        "ms": round(gmx.mean() * 1e3, 3), "bound": "mfma",
                        "achieved": round(ach_mx / 1e12, 2), "peak": MI355X_PEAK["bf16"] / 1e12,
                        "unit": "TFLOP/s", "frac": round(ach_mx / MI355X_PEAK["bf16"], 4),
                        "traffic": traffic.get("sketch_blk_kernel"),
                        "algorithmic": "2*m*n*N*l = %.4e FLOP per launch" % flops_bf16},
        "gram_gemm": {"kernel": "gemm_f64_dma_kernel (G = Y'Y, FP64 MFMA, lower-triangle tiles)",
                      "ms": round(grmx.mean() * 1e3, 3), "bound": "mfma",
                      "achieved": round(flops_gram / grmx.mean() / 1e12, 2),
                      "peak": MI355X_PEAK["f64"] / 1e12, "unit": "TFLOP/s",
                      "frac": round(flops_gram / grmx.mean() / MI355X_PEAK["f64"], 4),
                      "traffic": traffic.get("gemm_f64_dma_kernel_gram"),
                      "algorithmic": "m*Nl*(Nl+1) = %.4e FLOP per launch" % flops_gram},
    }
    dominant = max(kernels, key=lambda key: kernels[key]["ms"])
    value = args.steps / dtmx
    fp64_value = fp64_steps / dt64
    out = {
        "metric": "iSVD solves/s 100k×10k N=64: mixed vs FP64 at σ err≤1e-5; serve p50/p99",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dtmx / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 sketch (fp32 acc), f64 gram/eig/kn/svd",
        "data": "synthetic: A = U50 diag(0.9^i) V50' + 1e-3 noise floor, random Haar factors; "
                "Omega Philox seed 2",
        "config": {"workload": "iSVD 100000x10000 k=20 p=12 N=64, KN tol 1e-4, SVD former; "
                               "mixed (A FP32) vs plain FP64 (A FP64) on the same Omega",
                   "m": m, "n": n, "k": k, "p": p, "N": N, "precision": args.mixed},
        "fp64": {"value": round(fp64_value, 3), "unit": "solves/s",
                 "ms_per_step": round(dt64 / fp64_steps * 1e3, 3), "steps": fp64_steps,
                 "iterations": s64.iteration(),
                 "stage_ms": [round(x * 1e3, 3) for x in st64.mean(axis=0)],
                 "sketch_gemm_ms": round(g64.mean() * 1e3, 3),
                 "gram_gemm_ms": round(k64[:, 1].mean() * 1e3, 3),
                 "sketch_gemm_tflops": round(ach_64 / 1e12, 2),
                 "sketch_gemm_frac_of_f64_peak": round(ach_64 / MI355X_PEAK["f64"], 3)},
        "mixed": {"iterations": smx.iteration(),
                  "stage_ms": [round(x * 1e3, 3) for x in stmx.mean(axis=0)],
                  "sigma_rel_err_vs_fp64": rel, "subspace_dist_vs_fp64": dist,
                  "speedup_vs_fp64": round(value / fp64_value, 2),
                  "kernels": kernels},
        "stage_order": ["sketch", "orthogonalize", "integrate", "form"],
        "roofline": {k: v for k, v in kernels[dominant].items() if k != "ms"},
    }
    if not args.no_cpu:
        log("bench: cpu baseline (oracle, %d-row sample)" % args.cpu_rows)
        out["cpu_baseline"] = cpu_baseline(m, n, k, p, N, s64.iteration(), args.cpu_rows)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
