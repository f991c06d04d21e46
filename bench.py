"""iSVD solves/s on one MI355X: A 100000 x 10000, k=20, p=12, N=64 (Y is 100000 x 2048), KN tol
1e-4, SVD former.  The plain FP64 path (A FP64) and the mixed path (the same A stored FP32, BF16
MFMA sketch with FP32 accumulation, FP64 Gram/eig/KN/SVD) run on the same Philox Omega; `value`
is the mixed path's solves/s, reported beside the FP64 path's and the sigma relative error
between them.  The FP16 modes (one product with per-row scaling; two products of a hi/lo split of
A against an Omega that is exactly FP16) are reported beside it, each against an FP64 solve on
its own Omega.  Prints one JSON line.

A is synthetic: A = U50 diag(0.9^i) V50' + E with Haar U50 (m x 50), V50 (n x 50) and
E_ij = 1e-3 g_ij / (sqrt(m) + sqrt(n)), g ~ N(0,1) (||E||_2 ~ 1e-3); it is resident in HBM
before timing starts.

The `serve` object is the serving loop on the mixed-size request stream (Poisson arrivals at 75% of
the capacity measured by a closed-loop burst): solves/s and p50/p99 latency, submit to results.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK = {"bf16": 2.5e15, "f16": 2.5e15, "f64": 78.6e12, "hbm": 8e12}

# mixed mode -> (precision name, Omega mode name, sketch products, 16-bit format)
MIXED = {"bf16": ("PREC_MIXED_BF16", "OMEGA_PHILOX", 1, "bf16"),
         "bf16x3": ("PREC_MIXED_BF16X3", "OMEGA_PHILOX", 3, "bf16"),
         "f16": ("PREC_MIXED_F16", "OMEGA_PHILOX", 1, "f16"),
         "f16x2": ("PREC_MIXED_F16X2", "OMEGA_PHILOX_F16", 2, "f16"),
         "f16x3": ("PREC_MIXED_F16X3", "OMEGA_PHILOX", 3, "f16")}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_factors(m, n, r, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return U, V


def timed_solves(mc, solver, A, steps, warmup):
    """Times `steps` solves (after `warmup`); also returns, per solve, the HIP-event durations of
    the sketch GEMM, Gram GEMM and former Z = A'Qbar launches and of the four stages."""
    for _ in range(warmup):
        solver.solve(A)
    mc.synchronize()
    kern, stages = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.solve(A)
        times, _ = solver.stage_times()
        kern.append(solver.kernel_times() + (solver.kernel_time(mc._lib.KERNEL_FORMER_Z),))
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


def cpu_threads():
    """Threads the host BLAS runs with (and the CPUs this process may use)."""
    try:
        import threadpoolctl
        info = threadpoolctl.threadpool_info()
        blas = [d["num_threads"] for d in info if d.get("user_api") == "blas"]
        if blas:
            return int(blas[0])
    except Exception:
        pass
    return len(os.sched_getaffinity(0))


def cpu_baseline(m, n, k, p, N, rows):
    """The reference's CPU pipeline on the host cores: every m-sized product through numpy's
    multithreaded BLAS, the l x l and l x n factorizations through the oracle's small kernels
    (oracle/blas_pipeline.py, pinned against the reference's fixtures and the C oracle).  It runs
    the same workload family at `rows` of the m rows (A = U50 diag(0.9^i) V50' + noise floor, same
    n, k, p, N, KN to convergence); every stage is linear in m apart from the l x l steps, so the
    solve time is scaled by m / rows."""
    from oracle import blas_pipeline
    ms = rows
    rng = np.random.default_rng(7)
    U, V = make_factors(ms, n, 50, 8)
    A = (U * 0.9 ** np.arange(50)) @ V.T + 1e-3 * rng.standard_normal((ms, n)) / (np.sqrt(ms) + np.sqrt(n))
    A = np.asfortranarray(A)
    om = rng.standard_normal((n, N * (k + p)))
    blas_pipeline.solve(A[:2000], om, k, p, N, maxiter=2)  # warm the BLAS threads
    t0 = time.perf_counter()
    S, U_, it, stages = blas_pipeline.solve(A, om, k, p, N)
    sample_s = time.perf_counter() - t0
    scale = m / ms
    total = sample_s * scale
    threads = cpu_threads()
    return {"value": 1.0 / total, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": "reference pipeline (sketch, Gramian orthogonalizer, row-block KN to "
                      "convergence: %d iterations, SVD former) with numpy/OpenBLAS GEMMs on %d "
                      "threads and the oracle's small kernels, at %d of %d rows (%.0f%%): %.2f s, "
                      "stages %s s, scaled by m" % (it, threads, ms, m, 100.0 * ms / m, sample_s,
                                                    [round(x, 3) for x in stages]),
            "seconds_per_solve": total, "stage_seconds_sample": [round(x, 4) for x in stages]}


class MatrixRing:
    """Device memory for the client's per-request matrices: regions are taken front to back
    (wrapping to the start) and each is given back when its request is released; a region that
    is still in use blocks the next allocation until it is released."""

    def __init__(self, mc, nbytes):
        import threading
        self.buf = mc.DeviceArray((nbytes // 4,), np.float32)
        self.cap = nbytes
        self.head = 0
        self.live = {}  # key -> (offset, size)
        self.cv = threading.Condition()

    def _free_at(self, off, size):
        return all(off + size <= o or o + s <= off for o, s in self.live.values())

    def take(self, size, key):
        size = (size + 255) // 256 * 256
        assert size <= self.cap
        with self.cv:
            while True:
                off = self.head if self.head + size <= self.cap else 0
                if self._free_at(off, size):
                    break
                self.cv.wait(1.0)
            self.live[key] = (off, size)
            self.head = off + size
        return self.buf.ptr + off

    def give(self, key):
        with self.cv:
            self.live.pop(key, None)
            self.cv.notify_all()

    def free(self):
        self.buf.free()


def serve_bench(mc, args):
    """Continuous-batching serving of the request stream (mcnla_amd.workloads.serving_stream):
    every request brings its own FP32 m x n matrix with sigma_i = rho^i, written into device
    memory by the client on its own stream ahead of the arrival clock; the request carries the
    event that completes when its matrix is written, and its solve waits for that event on the
    device (so a late matrix shows up in the latency, which runs from submit to results).
    Capacity from a closed-loop burst, then Poisson arrivals at 75% of it; p50/p99 latency from
    submit to results ready."""
    import threading
    from mcnla_amd import workloads
    reqs = workloads.serving_stream(args.serve_requests + args.serve_burst)
    burst, stream = reqs[:args.serve_burst], reqs[args.serve_burst:]
    kw = dict(hbm_budget=int(args.serve_budget_gb * (1 << 30)), precision=mc.PREC_MIXED_BF16)
    gen = workloads.ReflectedGenerator(50000, 5000)
    ring = MatrixRing(mc, int(args.serve_ring_gb * (1 << 30)))

    def make(r, key, pre=None):
        """The request's matrix, enqueued on the generator's stream: (A, event when written).
        pre: its (sigma, (u, v)) computed beforehand."""
        ptr = ring.take(r["m"] * r["n"] * 4, key)
        A = mc.DeviceArray.view(ptr, (r["m"], r["n"]), np.float32, "F")
        if pre is None:
            return A, gen.write_async(A, workloads.stream_sigma(r), r["a_seed"])
        return A, gen.write_async(A, pre[0], r["a_seed"], vectors=pre[1])

    def submit(srv, r, A_ev):
        A, ev = A_ev
        return srv.submit(A, r["m"], r["n"], r["k"], r["p"], r["N"], seed=r["omega_seed"],
                          ready_event=ev)

    # the same requests alone (one slot, one at a time), to compare with their results in the mix
    check_idx = list(range(0, min(64, len(stream)), 4))  # kept however the stream is cut
    alone = {}
    with mc.Server(slots=1, **kw) as srv:
        for i in check_idx:
            A = make(stream[i], ("alone", i))
            rid = submit(srv, stream[i], A)
            srv.wait(rid)
            r = srv.result(rid)
            alone[i] = (r.S.numpy(), r.Qbar.numpy(), r.iterations)
            srv.release(rid)
            ring.give(("alone", i))
    log("serve: %d requests solved alone for the comparison" % len(check_idx))

    with mc.Server(slots=args.serve_slots, batch_max=args.serve_batch, **kw) as srv:
        mats = [make(r, ("burst", i)) for i, r in enumerate(burst)]
        mc.lib().mcnla_stream_synchronize(gen.stream)  # every burst matrix written
        t0 = time.perf_counter()
        ids = [submit(srv, r, A) for r, A in zip(burst, mats)]
        for _ in ids:
            srv.release(srv.next_done())
        cap = len(burst) / (time.perf_counter() - t0)
        for i in range(len(burst)):
            ring.give(("burst", i))
        log("serve: capacity %.1f solves/s (closed-loop burst of %d)" % (cap, len(burst)))
        if not stream:
            st = srv.stats()
            gen.free()
            ring.free()
            return {"capacity_solves_per_s": round(cap, 2), "slots": args.serve_slots,
                    "batch_max": args.serve_batch, "admissions": st["batches"],
            "kn_group_mean": round(st["kn_group_mean"], 2),
            "kn_chunk_ms": round(st["kn_chunk_mean_s"] * 1e3, 3),
                    "max_running": st["max_running"]}
        rate = args.serve_rate if args.serve_rate > 0 else 0.75 * cap
        if len(stream) > rate * args.serve_seconds:  # keep the open-loop phase bounded
            stream = stream[:max(64, int(rate * args.serve_seconds))]
            log("serve: stream cut to %d requests (%.0f s at %.1f/s)"
                % (len(stream), args.serve_seconds, rate))
        srv.reset_stats()
        arrivals = np.cumsum(np.random.default_rng(4).exponential(1.0 / rate, len(stream)))
        matched, mismatched = [0], [0]
        id_to_idx = {}
        lock = threading.Lock()

        queue_s, run_s = [], []

        def collect():
            for _ in range(len(stream)):
                rid = srv.next_done()
                with lock:
                    i = id_to_idx.get(rid)
                r = srv.result(rid)
                queue_s.append(r.queue_s)
                run_s.append(r.run_s)
                if i in alone:
                    got = (r.S.numpy(), r.Qbar.numpy(), r.iterations)
                    ok = all(np.array_equal(x, y) for x, y in zip(got[:2], alone[i][:2]))
                    ok = ok and got[2] == alone[i][2]
                    (matched if ok else mismatched)[0] += 1
                srv.release(rid)
                ring.give(("stream", i))

        # A producer writes the request matrices ahead of the arrival clock (up to `ahead` of
        # them), so that a request's matrix is resident when it arrives; `late` counts arrivals
        # whose matrix was not ready yet (submitted as soon as it was).
        # the requests' sigma and random vectors, drawn before the arrival clock starts
        pre = [(workloads.stream_sigma(r), workloads.ReflectedGenerator.vectors(r["m"], r["n"],
                                                                              r["a_seed"]))
               for r in stream]
        import queue
        ready = queue.Queue(maxsize=args.serve_ahead)

        def produce():
            for i, r in enumerate(stream):
                ready.put((i, make(r, ("stream", i), pre[i])))

        th = threading.Thread(target=collect)
        pr = threading.Thread(target=produce)
        pr.start()
        while ready.qsize() < min(args.serve_ahead, len(stream)) and pr.is_alive():
            time.sleep(0.001)
        th.start()
        late = 0
        t0 = time.perf_counter()
        for i, r in enumerate(stream):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            if ready.empty():
                late += 1
            j, A = ready.get()  # (matrix, event): its solve waits on the device for the event
            assert j == i
            with lock:
                id_to_idx[submit(srv, r, A)] = i
            if i % 512 == 511:
                log("serve: %d/%d submitted" % (i + 1, len(stream)))
        submit_span = time.perf_counter() - t0
        pr.join()
        th.join()
        st = srv.stats()
    gen.free()
    ring.free()
    log("serve: p50 %.2f ms, p99 %.2f ms, %.1f solves/s, max running %d, %d late arrivals"
        % (st["p50_s"] * 1e3, st["p99_s"] * 1e3, st["completed"] / st["span_s"], st["max_running"],
           late))
    return {"requests": len(stream), "arrivals": "Poisson", "rate_per_s": round(rate, 2),
            "offered_per_s": round(len(stream) / submit_span, 2), "late_arrivals": late,
            "capacity_solves_per_s": round(cap, 2), "solves_per_s": round(st["completed"] / st["span_s"], 2),
            "p50_ms": round(st["p50_s"] * 1e3, 3), "p99_ms": round(st["p99_s"] * 1e3, 3),
            "queue_p50_ms": round(float(np.percentile(queue_s, 50)) * 1e3, 3),
            "run_p50_ms": round(float(np.percentile(run_s, 50)) * 1e3, 3),
            "run_p99_ms": round(float(np.percentile(run_s, 99)) * 1e3, 3),
            "mean_ms": round(st["mean_s"] * 1e3, 3), "max_ms": round(st["max_s"] * 1e3, 3),
            "completed": st["completed"], "failed": st["failed"], "slots": args.serve_slots,
            "batch_max": args.serve_batch, "admissions": st["batches"],
            "kn_group_mean": round(st["kn_group_mean"], 2),
            "kn_chunk_ms": round(st["kn_chunk_mean_s"] * 1e3, 3),
            "max_running": st["max_running"], "hbm_budget_gb": args.serve_budget_gb,
            "peak_gb": round(st["peak_bytes"] / 2 ** 30, 2), "precision": "bf16 sketch, f64 rest",
            "alone_vs_mix_identical": "%d/%d" % (matched[0], matched[0] + mismatched[0]),
            "workload": "m log-U[1000,50000], n log-U[500,5000] (n<=m), k U{5..50}, p=10, "
                        "N U{8..32}, rho U(0.8,0.98); each request's A is a fresh FP32 m x n "
                        "matrix (I - 2uu')[diag(rho^i); 0](I - 2vv') with singular values "
                        "exactly rho^i, written on the device before submission; KN tol 1e-4, "
                        "SVD former, stream seed 3"}


def host_stream_bench(mc, P, A64, A32, s64, smx, args):
    """The same c2/c3 solves on a host A in pinned memory (mcnla_solver_solve_host): A streamed
    over PCIe in chunks beside the sketch of the chunks already on the device, against the plain
    H2D copy of A alone and the device-resident solve.  PCIe-inclusive rates; not the headline."""
    import ctypes
    L = mc.lib()
    out = {}
    for name, dA, s, ctype in (("fp64", A64, s64, ctypes.c_double), ("mixed", A32, smx,
                                                                     ctypes.c_float)):
        m, n = dA.shape
        hp = ctypes.c_void_p()
        mc._lib.check(L.mcnla_host_malloc(ctypes.byref(hp), dA.nbytes))
        try:
            flat = np.ctypeslib.as_array(ctypes.cast(hp.value, ctypes.POINTER(ctype)),
                                         shape=(m * n,))
            hA = flat.reshape((m, n), order="F")
            mc._lib.check(L.mcnla_memcpy_dtoh(hp.value, dA.ptr, dA.nbytes))
            mc._lib.check(L.mcnla_memcpy_htod(dA.ptr, hp.value, dA.nbytes))  # pages mapped
            t0 = time.perf_counter()
            mc._lib.check(L.mcnla_memcpy_htod(dA.ptr, hp.value, dA.nbytes))
            h2d = time.perf_counter() - t0
            s.solve(dA)
            S_dev = s.vectorS()
            s.solve_host(hA)  # warm-up: the solver's device copy of A is allocated here
            t0 = time.perf_counter()
            for _ in range(args.host_steps):
                s.solve_host(hA)
            dt = (time.perf_counter() - t0) / args.host_steps
            S_host = s.vectorS()
            out[name] = {"ms_per_solve": round(dt * 1e3, 2), "solves_per_s": round(1.0 / dt, 3),
                         "h2d_copy_ms": round(h2d * 1e3, 2),
                         "h2d_GBps": round(dA.nbytes / h2d / 1e9, 1),
                         "sigma_rel_vs_device_solve": float(np.max(np.abs(S_host - S_dev) / S_dev)),
                         "a_bytes": int(dA.nbytes)}
            log("bench: host-streamed %s %.1f ms/solve (H2D of A alone %.1f ms, %.1f GB/s)"
                % (name, dt * 1e3, h2d * 1e3, dA.nbytes / h2d / 1e9))
            del flat, hA
        finally:
            L.mcnla_host_free(hp.value)
    return out


def c4_bench(mc, args):
    """The tall row-block workload with the Wen-Yin integrator: A 500000 x 2000 FP32 row-major
    with sigma_i = 1/(1+i)^2, k=50, p=14, N=128 (Nl = 8192 > n).  The solver's automatic route
    forms G = Omega'(A'A)Omega instead of the 500000 x 8192 collection; one solve on the sketch
    route (FP64 sketch and Gram) is the reference for the sigma error."""
    m, n, k, p, N = 500000, 2000, 50, 14, 128
    from mcnla_amd import workloads
    blocks = m // n
    sig = 1.0 / (1.0 + np.arange(n)) ** 2
    Q, _ = np.linalg.qr(np.random.default_rng(42).standard_normal((n, n)))
    V, _ = np.linalg.qr(np.random.default_rng(43).standard_normal((n, n)))
    M = ((Q * sig) @ V.T / np.sqrt(blocks)).astype(np.float32)
    host, _ = workloads.sign_block_stack(blocks, M, seed=44)
    A = mc.DeviceArray.from_numpy(host)
    del host
    sig = np.sqrt(blocks) * np.linalg.svd(M.astype(np.float64), compute_uv=False)  # exact
    log("c4: A %d x %d FP32 row-major resident (sigma_i = 1/(1+i)^2)" % (m, n))
    P = mc.Parameters().setSize(m, n).setRank(k).setOverRank(p).setNumSketch(N)
    kw = dict(sketcher=mc.RowBlockGaussianProjectionSketcher,
              integrator=mc.RowBlockWenYinIntegrator, former=mc.RowBlockGramianFormer, seed=2)
    # FP16 with per-row scaling as the mixed precision (on the Gram-of-A route it forms the basis
    # Q = A X of the FP32 A on FP32 MFMA; there is no 16-bit sketch there), and the integrator's
    # chunks of iterations replayed as HIP graphs recorded in the warm-up solve
    s = mc.Solver(P, precision=mc.PREC_MIXED_F16, graph=args.c4_graph, **kw)
    tuning = tune(s, A, args, "c4")
    dt, kern, stages = timed_solves(mc, s, A, args.c4_steps, 1)
    S = s.vectorS()
    it = s.iteration()
    log("c4: %.1f ms/solve (auto route), %d WY iterations" % (dt / args.c4_steps * 1e3, it))
    ref = mc.Solver(P, route=mc.ROUTE_SKETCH, **kw)
    dref, _, sref = timed_solves(mc, ref, A, 1, 0)
    Sr = ref.vectorS()
    log("c4: sketch route (FP64 sketch + Gram of Y) %.1f ms/solve, %d iterations"
        % (dref * 1e3, ref.iteration()))
    out = {"workload": "iSVD 500000x2000 FP32 row-major, sigma_i=1/(1+i)^2, k=50 p=14 N=128 "
                       "(Nl=8192), row-block Wen-Yin tol 1e-3, row-block Gramian former",
           "route": "gram_of_a (auto)", "precision": "MIXED_F16", "graph": bool(args.c4_graph),
           "solves_per_s": round(args.c4_steps / dt, 3),
           "ms_per_solve": round(dt / args.c4_steps * 1e3, 3), "steps": args.c4_steps,
           "iterations": it, "stage_ms": [round(x * 1e3, 3) for x in stages.mean(axis=0)],
           "tuning": tuning,
           "sketch_route_ms_per_solve": round(dref * 1e3, 3),
           "sketch_route_stage_ms": [round(x * 1e3, 3) for x in sref.mean(axis=0)],
           "sketch_route_iterations": ref.iteration(),
           "sigma_rel_err_vs_sketch_route": float(np.max(np.abs(S - Sr) / Sr)),
           "sigma_vs_generator_max_rel": float(np.max(np.abs(S - sig[:k]) / sig[:k])),
           "data": "A = [diag(s_t) M]_t, 250 row blocks of 2000 with random signs s_t, "
                   "M = Q diag(1/(1+i)^2) V' / sqrt(250) (Q, V Haar 2000x2000) in FP32: the "
                   "block stack has orthonormal columns, so sigma(A) is 1/(1+i)^2 up to the FP32 "
                   "rounding of M and is known exactly; Omega Philox seed 2"}
    del s, ref
    A.free()
    return out


def solo_requests(mc, args):
    """Serving-stream requests solved one at a time on one thread (no server), each with its own
    solver, after one untimed solve of its shape: per-request wall time, stage times and
    iterations (the latency floor of a request, and a profile target without concurrent
    threads)."""
    from mcnla_amd import workloads
    reqs = workloads.serving_stream(args.solo_requests)
    gen = workloads.ReflectedGenerator(50000, 5000)
    out = []
    for r in reqs:
        A = gen.make(r["m"], r["n"], workloads.stream_sigma(r), r["a_seed"])
        P = mc.Parameters().setSize(r["m"], r["n"]).setRank(r["k"]).setOverRank(r["p"]).setNumSketch(r["N"])
        s = mc.Solver(P, seed=r["omega_seed"], precision=mc.PREC_MIXED_BF16)
        s.solve(A)
        mc.synchronize()
        t0 = time.perf_counter()
        s.solve(A)
        mc.synchronize()
        dt = time.perf_counter() - t0
        st, _ = s.stage_times()
        out.append({"m": r["m"], "n": r["n"], "l": r["k"] + r["p"], "N": r["N"],
                    "iterations": s.iteration(), "ms": round(dt * 1e3, 3),
                    "stage_ms": [round(x * 1e3, 3) for x in st]})
        del s
        A.free()
    gen.free()
    ms = np.array([o["ms"] for o in out])
    log("solo: %d requests, mean %.2f ms, p50 %.2f ms, p99 %.2f ms, mean iterations %.1f"
        % (len(out), ms.mean(), np.percentile(ms, 50), np.percentile(ms, 99),
           np.mean([o["iterations"] for o in out])))
    return {"requests": len(out), "mean_ms": round(float(ms.mean()), 3),
            "p50_ms": round(float(np.percentile(ms, 50)), 3),
            "stage_ms_mean": [round(float(x), 3) for x in np.mean([o["stage_ms"] for o in out], axis=0)],
            "mean_iterations": round(float(np.mean([o["iterations"] for o in out])), 2),
            "per_request": out}


def tune(solver, A, args, name):
    """Runs the autotuner on this solver (untimed set-up) unless --autotune 0; returns its plan."""
    if args.autotune:
        t0 = time.perf_counter()
        solver.autotune(A, reps=2)
        log("%s: autotuned in %.1f s: %s" % (name, time.perf_counter() - t0, solver.tuning()))
    return solver.tuning()


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
    ap.add_argument("--cpu-rows", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed", default="bf16", choices=sorted(MIXED))
    ap.add_argument("--accurate", default="f16,f16x2", help="comma-separated extra mixed modes "
                    "reported beside the headline one, each against an FP64 solve on its Omega "
                    "('' for none)")
    ap.add_argument("--serve-requests", type=int, default=4096)
    ap.add_argument("--serve-burst", type=int, default=256)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-budget-gb", type=float, default=48.0)
    ap.add_argument("--serve-seconds", type=float, default=90.0)
    ap.add_argument("--serve-ring-gb", type=float, default=32.0,
                    help="device memory for the client's per-request matrices")
    ap.add_argument("--serve-ahead", type=int, default=64,
                    help="request matrices the client writes ahead of the arrival clock")
    ap.add_argument("--serve-rate", type=float, default=0.0, help="Poisson arrival rate of the "
                    "open-loop phase (requests/s); 0: 75%% of the measured capacity")
    ap.add_argument("--serve-batch", type=int, default=1, help="KN requests a serving worker "
                    "admits at once and solves as one batch (their iterations in lockstep)")
    ap.add_argument("--serve-group-slots", type=int, default=4, help="workers of the second "
                    "serving run, through KN groups at the first run's arrival rate (0: skip it)")
    ap.add_argument("--serve-group-max", type=int, default=16, help="members per KN group of the "
                    "second serving run")
    ap.add_argument("--host-steps", type=int, default=1, help="c2/c3 solves of a host A streamed "
                    "over PCIe (mcnla_solver_solve_host), per precision; 0 skips them")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--only-serve", action="store_true", help="run the serving workload alone "
                    "and print its record")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0)
    ap.add_argument("--c4-steps", type=int, default=3)
    ap.add_argument("--graph", type=int, default=0, help="replay the integrator iterations as "
                    "HIP graphs recorded in the warm-up solve")
    ap.add_argument("--c4-graph", type=int, default=1, help="replay c4's integrator iterations "
                    "as HIP graphs (recorded in its warm-up solve)")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--autotune", type=int, default=1, help="tune each solver's split-K plans "
                    "before its warm-up (untimed); 0 keeps the default plans")
    ap.add_argument("--solo-requests", type=int, default=0, help="solve this many serving-stream "
                    "requests one at a time (no server) and print their record alone")
    ap.add_argument("--only-c4", action="store_true", help="run the row-block Wen-Yin workload "
                    "alone and print its record (for profiling)")
    args = ap.parse_args()
    # The serving phase runs 2 streams per slot (a solve and its side stream; 3 with a KN group:
    # the group's iterations, the members' other stages and the side stream); HIP maps streams
    # onto GPU_MAX_HW_QUEUES hardware queues per process (4 by default), and streams that share a
    # queue run their kernels one after another.  Before the runtime starts, give the process one
    # queue per stream of the serving slots (at most 32).
    queues = max((3 if args.serve_batch > 1 else 2) * args.serve_slots,
                 3 * args.serve_group_slots)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, queues))
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the iSVD path here is built for one MI355X "
              "(multi-GPU sharding is a design note, not built); rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import mcnla_amd as mc
    if mc.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.calibrate_copy_mb:
        # a plain device-to-device copy of known size: the reference reading for the HBM
        # traffic counters (FETCH_SIZE / WRITE_SIZE) of the roofline's `traffic`
        nbytes = int(args.calibrate_copy_mb) << 20
        src = mc.DeviceArray((nbytes // 8,), np.float64)
        dst = mc.DeviceArray((nbytes // 8,), np.float64)
        for _ in range(3):
            mc.lib().mcnla_memcpy_dtod(dst.ptr, src.ptr, nbytes, None)
        mc.synchronize()
        log("bench: copied %d bytes device-to-device 3 times" % nbytes)
        return
    if args.only_c4:
        print(json.dumps({"c4": c4_bench(mc, args)}), flush=True)
        return
    if args.solo_requests:
        print(json.dumps({"solo": solo_requests(mc, args)}), flush=True)
        return
    if args.only_serve:
        print(json.dumps({"serve": serve_bench(mc, args)}), flush=True)
        return
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
    pname, oname, terms, fmt = MIXED[args.mixed]
    prec, omode = getattr(mc, pname), getattr(mc, oname)
    s64 = mc.Solver(P, seed=2, graph=args.graph, omega_mode=omode)
    smx = mc.Solver(P, seed=2, precision=prec, graph=args.graph, omega_mode=omode)

    tune64 = tune(s64, A64, args, "fp64")
    tunemx = tune(smx, A32, args, "mixed")
    fp64_steps = args.fp64_steps if args.fp64_steps is not None else max(3, args.steps // 2)
    dt64, k64, st64 = timed_solves(mc, s64, A64, fp64_steps, 1)
    g64 = k64[:, 0]
    log("bench: fp64 %.2f ms/solve (sketch gemm %.2f ms, gram gemm %.2f ms), iterations %d"
        % (dt64 / fp64_steps * 1e3, g64.mean() * 1e3, k64[:, 1].mean() * 1e3, s64.iteration()))
    dtmx, kmx, stmx = timed_solves(mc, smx, A32, args.steps, args.warmup)
    gmx, grmx = kmx[:, 0], kmx[:, 1]
    log("bench: mixed %s %.2f ms/solve (sketch gemm %.2f ms, gram gemm %.2f ms), iterations %d"
        % (args.mixed, dtmx / args.steps * 1e3, gmx.mean() * 1e3, grmx.mean() * 1e3,
           smx.iteration()))
    S64, Smx = s64.vectorS(), smx.vectorS()
    rel = float(np.max(np.abs(Smx - S64) / S64))
    Q64, Qmx = s64.matrixQbar(), smx.matrixQbar()
    # ||Q Q' - Qr Qr'||_2 = sin of the largest principal angle
    c = np.linalg.svd(Q64.T @ Qmx, compute_uv=False)
    dist = float(np.sqrt(max(0.0, 1.0 - c.min() ** 2)))

    # the other mixed modes, each beside an FP64 solve on the same Omega
    accurate = {}
    for name in [x for x in args.accurate.split(",") if x and x != args.mixed]:
        pn, on, _, _ = MIXED[name]
        ref = s64 if getattr(mc, on) == omode else mc.Solver(P, seed=2, omega_mode=getattr(mc, on))
        if ref is not s64:
            ref.solve(A64)
        sx = mc.Solver(P, seed=2, precision=getattr(mc, pn), omega_mode=getattr(mc, on))
        tune(sx, A32, args, name)
        dtx, kx, stx = timed_solves(mc, sx, A32, args.steps, args.warmup)
        Sr, Sx = ref.vectorS(), sx.vectorS()
        accurate[name] = {
            "value": round(args.steps / dtx, 3), "unit": "solves/s",
            "ms_per_step": round(dtx / args.steps * 1e3, 3),
            "omega": on, "iterations": sx.iteration(), "fp64_iterations": ref.iteration(),
            "sigma_rel_err_vs_fp64": float(np.max(np.abs(Sx - Sr) / Sr)),
            "speedup_vs_fp64": round((args.steps / dtx) / (fp64_steps / dt64), 2),
            "stage_ms": [round(x * 1e3, 3) for x in stx.mean(axis=0)],
            "sketch_gemm_ms": round(kx[:, 0].mean() * 1e3, 3)}
        log("bench: mixed %s %.2f ms/solve, sigma rel err %.2e vs fp64 on its Omega"
            % (name, dtx / args.steps * 1e3, accurate[name]["sigma_rel_err_vs_fp64"]))
        del sx, ref

    host = host_stream_bench(mc, P, A64, A32, s64, smx, args) if args.host_steps > 0 else None

    nl = N * l
    flops_mx = 2.0 * m * n * nl * terms
    flops_f64 = 2.0 * m * n * nl
    flops_gram = 1.0 * m * nl * (nl + 1)  # lower triangle with diagonal, 2 flops per product
    bytes_z = 4.0 * m * n + 8.0 * m * l  # the former reads A (FP32) and Qbar (FP64) once
    ach_mx = flops_mx / gmx.mean()
    ach_64 = flops_f64 / g64.mean()
    traffic = load_traffic()
    kernels = {
        "sketch_gemm": {"kernel": "sketch_blk_kernel (Y = A*Omega, %s MFMA x%d products, FP32 "
                                  "accumulate)" % (fmt.upper(), terms),
                        "ms": round(gmx.mean() * 1e3, 3), "bound": "mfma",
                        "achieved": round(ach_mx / 1e12, 2), "peak": MI355X_PEAK[fmt] / 1e12,
                        "unit": "TFLOP/s", "frac": round(ach_mx / MI355X_PEAK[fmt], 4),
                        "traffic": traffic.get("sketch_blk_kernel"),
                        "algorithmic": "2*m*n*N*l*products = %.4e FLOP per launch" % flops_mx},
        "former_z": {"kernel": "tall_stream_kernel (Z = A'Qbar, A FP32 and Qbar streamed into "
                               "LDS by DMA, FP32 MFMA)",
                     "ms": round(kmx[:, 2].mean() * 1e3, 3), "bound": "hbm",
                     "achieved": round(bytes_z / kmx[:, 2].mean() / 1e9, 1),
                     "peak": MI355X_PEAK["hbm"] / 1e9, "unit": "GB/s",
                     "frac": round(bytes_z / kmx[:, 2].mean() / MI355X_PEAK["hbm"], 4),
                     "traffic": traffic.get("tall_stream_kernel_former"),
                     "algorithmic": "4*m*n + 8*m*l = %.4e bytes per launch" % bytes_z},
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
        "dtype": "%s sketch (fp32 acc), f64 gram/eig/kn/svd" % fmt,
        "data": "synthetic: A = U50 diag(0.9^i) V50' + 1e-3 noise floor, random Haar factors; "
                "Omega %s seed 2" % oname,
        "config": {"workload": "iSVD 100000x10000 k=20 p=12 N=64, KN tol 1e-4, SVD former; "
                               "mixed (A FP32) vs plain FP64 (A FP64) on the same Omega",
                   "m": m, "n": n, "k": k, "p": p, "N": N, "precision": args.mixed},
        "fp64": {"value": round(fp64_value, 3), "unit": "solves/s",
                 "ms_per_step": round(dt64 / fp64_steps * 1e3, 3), "steps": fp64_steps,
                 "iterations": s64.iteration(), "tuning": tune64,
                 "stage_ms": [round(x * 1e3, 3) for x in st64.mean(axis=0)],
                 "sketch_gemm_ms": round(g64.mean() * 1e3, 3),
                 "gram_gemm_ms": round(k64[:, 1].mean() * 1e3, 3),
                 "former_z_ms": round(k64[:, 2].mean() * 1e3, 3),
                 "former_z_GBps": round((8.0 * m * n + 8.0 * m * l) / k64[:, 2].mean() / 1e9, 1),
                 "sketch_gemm_tflops": round(ach_64 / 1e12, 2),
                 "sketch_gemm_frac_of_f64_peak": round(ach_64 / MI355X_PEAK["f64"], 3)},
        "mixed": {"iterations": smx.iteration(), "tuning": tunemx,
                  "stage_ms": [round(x * 1e3, 3) for x in stmx.mean(axis=0)],
                  "sigma_rel_err_vs_fp64": rel, "subspace_dist_vs_fp64": dist,
                  "speedup_vs_fp64": round(value / fp64_value, 2),
                  "kernels": kernels},
        "mixed_modes": accurate,
        "host_stream": host,
        "stage_order": ["sketch", "orthogonalize", "integrate", "form"],
        "roofline": {k: v for k, v in kernels[dominant].items() if k != "ms"},
    }
    del s64, smx, A64, A32  # the other workloads' matrices and workspaces take their place
    if not args.no_c4:
        out["c4"] = c4_bench(mc, args)
    if not args.no_serve:
        out["serve"] = serve_bench(mc, args)
        if args.serve_group_slots > 0:
            # the same stream at the same arrival rate through KN groups (iteration-level batching)
            g = argparse.Namespace(**vars(args))
            g.serve_slots, g.serve_batch = args.serve_group_slots, args.serve_group_max
            g.serve_rate = out["serve"]["rate_per_s"]
            out["serve_grouped"] = serve_bench(mc, g)
    if not args.no_cpu:
        log("bench: cpu baseline (BLAS pipeline, %d-row sample)" % args.cpu_rows)
        out["cpu_baseline"] = cpu_baseline(m, n, k, p, N, min(args.cpu_rows, m))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
