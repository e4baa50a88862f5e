"""Throughput of the batched Newton solve on one MI355X: the learned start with the mixed-precision
method against the plain FP64 method.

Workload (the configuration the headline metric is quoted on): one implicit-midpoint step of the
Hénon–Heiles system for 2^20 independent states, n = 4 unknowns each, h = 0.1, states on energy
shells E ~ U[0.01, 0.15] (seed 0); f_abstol = 1e-12, f_reltol = 0, status form of solve_batch
(72-byte SolverStatus per system).

  plain          Newton + Static + static LU in FP64 from x₀ = x_k.
  learned+mixed  the learned start (a 3 × 128 BF16 MLP, trained before the clock starts on 10^6
                 solutions of the plain method, from states of another seed) followed by Newton with
                 LUFactorization(Float32) + Refinement, and plain Newton from x_k for every system
                 that does not converge from the learned start — all three in one fused kernel.

A "step" is one solve of the whole batch in the given form. value = converged systems/s of the
learned+mixed form: systems with code SUCCESS whose FP64 residual, rechecked on the host after the
timed region, satisfies ‖F(x)‖₂ ≤ 1e-12, divided by the wall time of the K timed steps; the plain
form's figure and the ratio sit beside it in `config`.

`--workload fpu_gl3`, `--workload expsum_lsq` and `--workload dense_sin` print the same kind of line for
the 65536-system FPU-β Gauss–Legendre stage systems (n = 48), the 262144 exponential-sum fits and the
2^20 dense systems of the reference's §10.E grid problem at n = 16. The default line also carries a
compact line of each under `config.sub_results`, each measured in a child process after the headline's
timed region (`--no-sub-results` leaves them out).
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

NSYS = 1 << 20
H = 0.1
TOL = 1e-12
NTRAIN = 1_000_000
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6
# algorithmic BF16 flops per system of the learned start: 2 × (5·128 + 128·128 + 128·128 + 128·4)
MLP_FLOPS_PER_SYSTEM = 2 * (5 * 128 + 128 * 128 + 128 * 128 + 128 * 4)
# algorithmic HBM bytes per system of the plain solve: x_k 32 (record) + X in 32 + X out 32 + S 72
BYTES_PER_SYSTEM = 32 + 32 + 32 + 72


def cpu_baseline(gs, xk, min_seconds: float = 8.0):
    """The oracle (C++ restatement of the reference's R1 loop, OpenMP over systems) on the host cores:
    the plain method on a bounded sample of the same workload, repeated until `min_seconds`."""
    from oracle import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    sample = xk[: 1 << 18]
    opts = gs.Options(f_abstol=TOL, f_reltol=0.0)
    orc.solve_batch(gs.Newton(), gs.HenonHeilesMidpoint(), sample.copy(), gs.BatchParams(H, sample), opts,
                    nthreads=threads)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        x, S = orc.solve_batch(gs.Newton(), gs.HenonHeilesMidpoint(), sample.copy(), gs.BatchParams(H, sample),
                               opts, nthreads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds:
            break
    conv = int((S["code"] == 0).sum())
    # the same port on one thread, on a smaller sample
    one = xk[: 1 << 14]
    reps1, t1 = 0, time.perf_counter()
    while True:
        _, S1 = orc.solve_batch(gs.Newton(), gs.HenonHeilesMidpoint(), one.copy(), gs.BatchParams(H, one), opts,
                                nthreads=1)
        reps1 += 1
        el1 = time.perf_counter() - t1
        if el1 >= min_seconds / 4:
            break
    return {
        "value": conv * reps / el,
        "unit": "converged systems/s",
        "cores": threads,
        "kind": "port",
        "sample": f"plain method on {len(sample)} systems of the same workload (seed 0) x {reps} repetitions, "
                  f"{el:.1f} s wall, OpenMP schedule(static) over systems",
        "single_thread": {"value": int((S1["code"] == 0).sum()) * reps1 / el1, "cores": 1,
                          "sample": f"{len(one)} systems x {reps1} repetitions, {el1:.1f} s wall"},
    }


def _event_ms(events):
    import torch

    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in events]


def _traffic(name):
    """HBM bytes per launch of a kernel, from the calibrated FETCH_SIZE + WRITE_SIZE counters
    committed under profiles/ (None where no such file exists)."""
    tfile = ROOT / "profiles" / name
    return json.loads(tfile.read_text()).get("bytes_per_launch") if tfile.exists() else None


def bench_henon_heiles(args, gs, workloads, torch):
    dev = torch.device("cuda:0")
    lib = gs.load_library()
    K, W = args.steps, args.warmup
    xk = workloads.henon_heiles_states(NSYS, seed=0)
    Xk = torch.as_tensor(xk, device=dev)
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    opts = gs.Options(f_abstol=TOL, f_reltol=0.0)
    p = gs.BatchParams(H, Xk)
    plain = gs.Newton()
    mixed = gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())

    # the network, trained before any clock starts
    net = gs.LearnedStart(dev, seed=0)
    if args.load_net:
        # a network this script trained and saved before (--save-net), so that a profiled run need
        # not train
        z = np.load(args.load_net)
        net.load_state({"params": z["params"], "norm": json.loads(str(z["norm"]))})
        training = {"loaded_from": args.load_net}
    else:
        print("bench: training the network", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        Xt = torch.as_tensor(workloads.henon_heiles_states(NTRAIN, seed=100), device=dev)
        X1, _ = gs.generate_solutions(gs, Xt, H, opts)
        net.fit_normalization(Xt, X1, H)
        hist = net.train(Xt, X1, H, epochs=args.epochs, batch=4096, lr=3e-3, lr_final=3e-5, seed=0)
        torch.cuda.synchronize()
        training = {"pairs": NTRAIN, "epochs": args.epochs, "seconds": time.perf_counter() - t0,
                    "final_loss": hist[-1]}
        del Xt, X1
        if args.save_net:
            st = net.state()
            np.savez(args.save_net, params=st["params"], norm=json.dumps(st["norm"]))

    # plain: K steps on fresh copies made before the clock
    print("bench: timed steps", file=sys.stderr, flush=True)
    starts = [Xk.clone() for _ in range(K + W)]
    Sp = [gs.status_buffer(NSYS, dev) for _ in range(K)]
    Sw = gs.status_buffer(NSYS, dev)
    lib.gs_set_kernel_timing(1)
    for i in range(W):
        gs.solve_batch(starts[K + i], Sw, prob, plain, p, opts)
    torch.cuda.synchronize()
    solve_ms = []
    t0 = time.perf_counter()
    for i in range(K):
        gs.solve_batch(starts[i], Sp[i], prob, plain, p, opts)
        solve_ms.append(lib.gs_last_kernel_ms())
    torch.cuda.synchronize()
    plain_el = time.perf_counter() - t0
    lib.gs_set_kernel_timing(0)

    # learned + mixed: one fused kernel per step (start, mixed solve from it, plain fallback); the
    # start overwrites X, so no copies are needed
    Xl = [torch.empty_like(Xk) for _ in range(K)]
    Sl = [gs.status_buffer(NSYS, dev) for _ in range(K)]
    Xw = torch.empty_like(Xk)
    for i in range(W):
        gs.solve_batch_learned(gs, Xw, Sw, prob, mixed, p, net, opts)
    torch.cuda.synchronize()
    lib.gs_set_kernel_timing(1)
    fused_ms = []
    fallbacks = 0
    t0 = time.perf_counter()
    for i in range(K):
        fallbacks += gs.solve_batch_learned(gs, Xl[i], Sl[i], prob, mixed, p, net, opts)
        fused_ms.append(lib.gs_last_kernel_ms())
    torch.cuda.synchronize()
    learned_el = time.perf_counter() - t0
    lib.gs_set_kernel_timing(0)
    fused_kms = float(np.mean(fused_ms))
    # the start kernel alone (not part of the timed step), for the network's share of the fused kernel:
    # 20 launches back to back between two events on the launch stream (the network's C record is built
    # once, so that the host keeps ahead of the GPU)
    import ctypes

    net_c = net.to_c()
    stream = gs._stream_ptr(Xk)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        rc = lib.gs_mlp_start_henon_heiles(ctypes.byref(net_c), NSYS, ctypes.c_void_p(Xk.data_ptr()), H,
                                           ctypes.c_void_p(Xw.data_ptr()), stream)
        assert rc == 0
    e1.record()
    torch.cuda.synchronize()
    start_ms = e0.elapsed_time(e1) / 20

    # host FP64 recheck of every step's result, outside the timed regions
    print("bench: host recheck and CPU baseline", file=sys.stderr, flush=True)
    def converged(Xs, Ss):
        n = 0
        for X, S in zip(Xs, Ss):
            st = gs.statuses(S)
            x = X.cpu().numpy()
            r = np.sqrt((workloads.henon_heiles_midpoint_residual(x, xk, H) ** 2).sum(axis=1))
            n += int(((st["code"] == 0) & (r <= TOL)).sum())
        return n

    conv_plain = converged(starts[:K], Sp)
    conv_learned = converged(Xl, Sl)
    it_plain = float(gs.statuses(Sp[-1])["iterations"].mean())
    it_learned = float(gs.statuses(Sl[-1])["iterations"].mean())
    value = conv_learned / learned_el
    plain_value = conv_plain / plain_el
    achieved = MLP_FLOPS_PER_SYSTEM * NSYS / (fused_kms * 1e-3) / 1e12
    traffic = _traffic("traffic_learned_solve_henon_heiles_2p20.json")
    solve_kms = float(np.mean(solve_ms))
    return {
        "metric": "converged systems/s at residual ≤1e-12 (FP64 check), learned+mixed vs plain, 1 GPU",
        "value": value,
        "unit": "converged systems/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": learned_el / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (numpy PCG64, Hénon–Heiles energy shells; batch seed 0, training seed 100)",
        "config": {
            "workload": "henon_heiles_implicit_midpoint n=4, batch 2^20, h=0.1, abstol 1e-12, reltol 0, status form",
            "method": "learned start (3x128 BF16 MLP) + Newton(LUFactorization(Float32), Refinement), "
                      "plain Newton fallback, fused in one kernel",
            "batch": NSYS,
            "learned_mixed_value": value,
            "plain_value": plain_value,
            "learned_mixed_vs_plain": value / plain_value,
            "plain_ms_per_step": plain_el / K * 1e3,
            "plain_solve_kernel_ms": solve_kms,
            "plain_solve_kernel_hbm_frac": BYTES_PER_SYSTEM * NSYS / (solve_kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "learned_solve_kernel_ms": fused_kms,
            "learned_start_kernel_alone_ms": start_ms,
            "mean_iterations_plain": it_plain,
            "mean_iterations_learned_mixed": it_learned,
            "fallback_systems": fallbacks,
            "training": training,
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "learned_solve_kernel (BF16 MFMA start + mixed Newton + fallback, fused); "
                      "achieved counts the network's BF16 flops only",
            "achieved": achieved,
            "peak": BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / BF16_PEAK_TFLOPS,
            "traffic": traffic,
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(gs, xk),
    }


# algorithmic flops of one Newton iteration of the n = 48 stage system: the LU (2/3 n³), the two
# triangular solves of the direction and one refinement pass (3 × 2n²), and the residual and Jacobian
# products (≈ 3000)
FPU_FLOPS_PER_ITER = 2 * 48 ** 3 // 3 + 6 * 48 * 48 + 3000
# algorithmic HBM bytes per system of the stage systems: (q, p) 256 + Z in 384 + Z out 384 + S 72
FPU_BYTES_PER_SYSTEM = 256 + 384 + 384 + 72


def _timed_oracle(fn, min_seconds):
    """Repeat fn() until min_seconds of wall time; return (repetitions, seconds, last result)."""
    fn()  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        out = fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds:
            return reps, el, out


def bench_fpu_gl3(args, gs, workloads, torch):
    """65536 FPU-β Gauss–Legendre s = 3 stage systems (n = 48): plain FP64 and mixed FP32 + refinement."""
    dev = torch.device("cuda:0")
    lib = gs.load_library()
    n = 65536
    qp = workloads.fpu_states(n, seed=0)
    QP = torch.as_tensor(qp, device=dev)
    prob = gs.NonlinearProblem(gs.FpuGl3())
    p = gs.BatchParams(0.05, QP)
    opts = gs.Options(f_abstol=TOL, f_reltol=0.0)
    out = {}
    for name, m in (("plain", gs.Newton()),
                    ("mixed", gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement()))):
        Zs = [torch.zeros((n, 48), dtype=torch.float64, device=dev) for _ in range(args.steps + args.warmup)]
        S = [gs.status_buffer(n, dev) for _ in range(args.steps)]
        Sw = gs.status_buffer(n, dev)
        for i in range(args.warmup):
            gs.solve_batch(Zs[args.steps + i], Sw, prob, m, p, opts)
        torch.cuda.synchronize()
        lib.gs_set_kernel_timing(1)
        kms = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            gs.solve_batch(Zs[i], S[i], prob, m, p, opts)
            kms.append(lib.gs_last_kernel_ms())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        lib.gs_set_kernel_timing(0)
        conv, iters = 0, 0
        for i in range(args.steps):
            st = gs.statuses(S[i])
            r = np.sqrt((workloads.fpu_gl3_residual(Zs[i].cpu().numpy(), qp, 0.05) ** 2).sum(axis=1))
            conv += int(((st["code"] == 0) & (r <= TOL)).sum())
            iters += int(st["iterations"].sum())
        out[name] = (conv / el, el / args.steps * 1e3, float(gs.statuses(S[-1])["iterations"].mean()),
                     float(np.mean(kms)), iters / args.steps)
    kms, iters = out["mixed"][3], out["mixed"][4]
    achieved = FPU_FLOPS_PER_ITER * iters / (kms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        k = 4096
        reps, cel, (_, so) = _timed_oracle(
            lambda: orc.solve_batch(gs.Newton(), gs.FpuGl3(), np.zeros((k, 48)), gs.BatchParams(0.05, qp[:k]), opts,
                                    nthreads=threads), 8.0)
        cpu = {"value": int((so["code"] == 0).sum()) * reps / cel, "unit": "converged systems/s", "cores": threads,
               "kind": "port", "sample": f"plain FP64 Newton on the first {k} systems of the same batch x {reps} "
                                         f"repetitions, {cel:.1f} s wall, OpenMP schedule(static) over systems"}
    return {
        "metric": "converged systems/s at residual ≤1e-12 (FP64 check), 1 GPU",
        "value": out["mixed"][0],
        "unit": "converged systems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["mixed"][1],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic ((q, p) ~ N(0, 0.3^2), numpy PCG64 seed 0)",
        "config": {"workload": "fpu_beta_gl3 n=48, batch 65536, h=0.05, beta=1, abstol 1e-12, reltol 0",
                   "method": "Newton(LUFactorization(Float32), Refinement); the factor's rows in VGPRs",
                   "plain_value": out["plain"][0], "plain_ms_per_step": out["plain"][1],
                   "mean_iterations": out["mixed"][2], "mixed_kernel_ms": out["mixed"][3],
                   "plain_kernel_ms": out["plain"][3]},
        "roofline": {
            "bound": "fp64 valu",
            "kernel": "fpu_gl3_kernel<mixed, refine> (one wave per system, each lane's factor row in VGPRs; "
                      "FP32 LU, FP64 refinement; VALU-bound, no MFMA); achieved = "
                      f"{FPU_FLOPS_PER_ITER} flop per Newton iteration x the batch's iterations / kernel time; "
                      "peak = the FP64 dense peak (equal to the FP64 vector rate), the precision of the residual, "
                      "refinement and update",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": _traffic("traffic_fpu_gl3_65536.json"),
            "hbm_frac": FPU_BYTES_PER_SYSTEM * n / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "cpu_baseline": cpu,
    }


def bench_expsum(args, gs, workloads, torch):
    """262144 exponential-sum least-squares fits (K ~ U{1..4}, m ~ U{16..128}), LevenbergMarquardt."""
    dev = torch.device("cuda:0")
    n = 262144
    K_, m_, off, Y, X0, _ = workloads.expsum_fits(n, seed=0)
    data = gs.ExpSumData(*(torch.as_tensor(v, device=dev) for v in (K_, m_, off, Y)))
    prob = gs.LeastSquaresProblem(gs.ExpSumFits())
    method = gs.Newton(globalization=gs.LevenbergMarquardt())
    opts = gs.Options()
    X0d = torch.as_tensor(X0, device=dev)
    # the launch shapes of the four buckets from the package's autotuner, on this batch
    t0 = time.perf_counter()
    tuned = gs.tune_lsq(X0d, data, method, opts, repeats=1)
    tune_s = time.perf_counter() - t0
    Xs = [X0d.clone() for _ in range(args.steps + args.warmup)]
    S = [gs.status_buffer(n, dev) for _ in range(args.steps)]
    Sw = gs.status_buffer(n, dev)
    for i in range(args.warmup):
        gs.solve_batch(Xs[args.steps + i], Sw, prob, method, data, opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        gs.solve_batch(Xs[i], S[i], prob, method, data, opts)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = gs.statuses(S[-1])
    codes = np.bincount(st["code"], minlength=6)
    # algorithmic HBM bytes of one step: y (8m), K and m (4 + 4), the offset (8), X in and out
    # (8n each) and the status (72) of every fit
    nbytes = float((8 * m_.astype(np.int64) + 16 + 2 * 8 * 2 * K_.astype(np.int64) + 72).sum())
    hbm_achieved = nbytes / (el / args.steps) / 1e9
    # algorithmic FP64 flops of the step: per LM iteration of a fit with K terms (n = 2K) and m points,
    # m (40K + 2n + 3n²) — the residual and Jacobian rows with their exponentials (≈ 40 per term), the
    # Givens reduction of [J | −F] (3n² per row) and the right-hand side (2n) — over each fit's own
    # iteration count (the damping trials beyond the first are not counted: a lower bound)
    Kf = K_.astype(np.float64)
    flops = float((st["iterations"].astype(np.float64) * m_.astype(np.float64) * (40 * Kf + 4 * Kf + 12 * Kf * Kf)).sum())
    achieved = flops / (el / args.steps) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        k = 8192
        reps, cel, _ = _timed_oracle(
            lambda: orc.solve_expsum_batch(method, opts, X0[:k], K_[:k], m_[:k], off[:k], Y, nthreads=threads), 8.0)
        cpu = {"value": k * reps / cel, "unit": "fits/s", "cores": threads, "kind": "port",
               "sample": f"the first {k} fits of the same batch x {reps} repetitions, {cel:.1f} s wall, "
                         "OpenMP schedule(static) over fits"}
    return {
        "metric": "fits/s (every fit run to its LevenbergMarquardt exit), 1 GPU",
        "value": n * args.steps / el,
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (numpy PCG64 seed 0)",
        "config": {"workload": "expsum_lsq 262144 fits, K~U{1..4}, m~U{16..128}, sigma 1e-3",
                   "method": "Newton + LevenbergMarquardt (stacked Givens), defaults",
                   "codes": codes.tolist(), "total_iterations": int(st["iterations"].sum()),
                   "launch": "device bucketing + four persistent bucket kernels, one captured hipGraph per call shape",
                   "tuned_shapes": {str(k): [v[0], v[1]] for k, v in tuned.items()}, "tuning_seconds": tune_s},
        "roofline": {
            "bound": "fp64 valu",
            "kernel": "lsq_kernel<2,4,6,8> (concurrent buckets); achieved = m (40K + 2n + 3n^2) flop per LM "
                      "iteration of each fit x its iterations / step time (a lower bound: damping trials not "
                      "counted); peak = the FP64 vector rate",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": _traffic("traffic_lsq_expsum_262144.json"),
            "hbm_achieved_gbs": hbm_achieved,
            "hbm_frac": hbm_achieved / HBM_PEAK_GBS,
        },
        "cpu_baseline": cpu,
    }


# algorithmic flops of one Newton iteration of the dense §10.E family at size n: the LU (2/3 n³), the
# two triangular solves (2n²) and the residual's matrix-vector product (2n²)
def dense_flops_per_iter(n: int) -> int:
    return 2 * n ** 3 // 3 + 4 * n * n


def bench_dense(args, gs, workloads, torch):
    """The §10.E problem of the reference's device grid (A x + sin.(x)/10 − b, benchmark/device_problems.jl:
    133-175) at n = args.dense_n (16 … 64) from x₀ = 0, plain FP64 Newton, status form: a wave segment
    per system, its width (lanes per system) from the package's autotuner on this batch before the
    clock. 2^20 systems up to n = 32, 2^18 above."""
    n = args.dense_n
    nsys = 1 << 20 if n <= 32 else 1 << 18
    dev = torch.device("cuda:0")
    lib = gs.load_library()
    A, B, _ = workloads.tene_batch(n, nsys)
    fam = gs.DenseSin(A)
    prob = gs.NonlinearProblem(fam)
    Bd = torch.as_tensor(B, device=dev)
    p = gs.PerSystem(Bd)
    opts = gs.Options(f_abstol=TOL, f_reltol=0.0)
    m = gs.Newton()
    tuned = gs.tune_dense(torch.zeros((nsys, n), dtype=torch.float64, device=dev), prob, m, p, opts, repeats=1)
    Xs = [torch.zeros((nsys, n), dtype=torch.float64, device=dev) for _ in range(args.steps + args.warmup)]
    S = [gs.status_buffer(nsys, dev) for _ in range(args.steps)]
    Sw = gs.status_buffer(nsys, dev)
    for i in range(args.warmup):
        gs.solve_batch(Xs[args.steps + i], Sw, prob, m, p, opts)
    torch.cuda.synchronize()
    lib.gs_set_kernel_timing(1)
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        gs.solve_batch(Xs[i], S[i], prob, m, p, opts)
        kms.append(lib.gs_last_kernel_ms())
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.gs_set_kernel_timing(0)
    conv, iters = 0, 0
    for i in range(args.steps):
        st = gs.statuses(S[i])
        r = np.sqrt((workloads.tene_residual(Xs[i].cpu().numpy(), A, B) ** 2).sum(axis=1))
        conv += int(((st["code"] == 0) & (r <= TOL)).sum())
        iters += int(st["iterations"].sum())
    kmean = float(np.mean(kms))
    achieved = dense_flops_per_iter(n) * iters / args.steps / (kmean * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        k = 1 << 15 if n <= 32 else 1 << 12
        reps, cel, (_, so) = _timed_oracle(
            lambda: orc.solve_batch(m, fam, np.zeros((k, n)), gs.PerSystem(B[:k]), opts, nthreads=threads), 8.0)
        cpu = {"value": int((so["code"] == 0).sum()) * reps / cel, "unit": "converged systems/s", "cores": threads,
               "kind": "port", "sample": f"plain FP64 Newton on the first {k} systems of the same batch x {reps} "
                                         f"repetitions, {cel:.1f} s wall, OpenMP schedule(static) over systems"}
    return {
        "metric": "converged systems/s at residual ≤1e-12 (FP64 check), 1 GPU",
        "value": conv / el,
        "unit": "converged systems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the §10.E batch: A = 2I + R from seed n, x* ~ U[-1,1]^n from seed 1, numpy PCG64)",
        "config": {"workload": f"dense_sin (§10.E A x + sin(x)/10 - b) n={n}, batch {nsys}, x0 = 0, abstol 1e-12, "
                               "reltol 0, status form",
                   "method": "Newton() + Static + static LU in FP64; a wave segment of 16, 32 or 64 lanes per system",
                   "lanes_per_system": tuned["lanes"],
                   "tuner_seconds": {str(k): v for k, v in tuned["seconds"].items()},
                   "kernel_ms": kmean, "mean_iterations": iters / args.steps / nsys,
                   "reference_published_n16": "2.11e7 systems/s (RTX 4090, Float32, gs-newton; benchmark/REPORT.md), "
                                              "context only"},
        "roofline": {
            "bound": "fp64 valu",
            "kernel": f"dense_kernel<{n}, {tuned['lanes']} lanes, Static, FP64, status>; achieved = (2/3 n^3 + 4 n^2) flop per Newton "
                      "iteration x the batch's iterations / kernel time",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }


FP32_PEAK_TFLOPS = 157.3

# The reference's published R2 grid cells for the §10.E problem in Float32 (benchmark/REPORT.md, the
# grid table): systems/s at 10^6 and 10^7 systems, per machine, n and variant. The ROCm card's band
# stops at n = 8 (benchmark/devices.jl BANDS).
def _cells(rows):
    return {v: {"1e6": a, "1e7": b} for v, a, b in rows}


REFERENCE_GRID = {
    "rtx4090": {
        2: _cells([("gs-newton", 2.10e9, 2.24e9), ("gs-reuse4", 6.61e8, 7.09e8), ("gs-backtracking", 1.55e9, 1.80e9)]),
        4: _cells([("gs-newton", 9.06e8, 1.01e9), ("gs-reuse4", 2.19e8, 2.29e8), ("gs-backtracking", 6.71e8, 7.24e8)]),
        6: _cells([("gs-newton", 2.93e8, 3.05e8), ("gs-reuse4", 9.69e7, 1.01e8), ("gs-backtracking", 2.31e8, 2.39e8)]),
        8: _cells([("gs-newton", 1.43e8, 1.47e8), ("gs-reuse4", 5.07e7, 5.19e7), ("gs-backtracking", 1.16e8, 1.20e8)]),
        12: _cells([("gs-newton", 4.49e7, 4.55e7), ("gs-reuse4", 2.01e7, 2.03e7), ("gs-backtracking", 3.72e7, 3.76e7)]),
        16: _cells([("gs-newton", 2.06e7, 2.11e7), ("gs-reuse4", 9.95e6, 1.01e7), ("gs-backtracking", 1.54e7, 1.56e7)]),
    },
    "rx7900xtx": {
        2: _cells([("gs-newton", 2.30e9, 2.05e9), ("gs-reuse4", 4.47e8, 7.11e8), ("gs-backtracking", 3.90e9, 2.32e9)]),
        4: _cells([("gs-newton", 2.49e8, 4.58e8), ("gs-reuse4", 1.56e8, 1.68e8), ("gs-backtracking", 2.07e8, 3.74e8)]),
        6: _cells([("gs-newton", 1.45e8, 2.72e8), ("gs-reuse4", 8.62e7, 9.59e7), ("gs-backtracking", 1.52e8, 1.98e8)]),
        8: _cells([("gs-newton", 6.05e7, 6.23e7), ("gs-reuse4", 1.98e7, 2.08e7), ("gs-backtracking", 4.45e7, 4.61e7)]),
    },
}


def bench_dense_grid(args, gs, workloads, torch):
    """The reference's device throughput grid (benchmark/devices.jl:446-500) at n = args.dense_n in its
    element type, Float32: the §10.E batch from x₀ = 0 with f_abstol = √eps(Float32), f_reltol = 0,
    max_iterations = 200, for the variants gs-newton (Newton()), gs-reuse4 (Newton(reuse = 4)) and
    gs-backtracking (Newton(Backtracking())), at 10^6 and 10^7 systems. As the reference's cell, a
    launch is timed to its synchronisation on a fresh start made before the clock (one warm-up, then
    the median of `steps` launches), and the cell is n_sys / that time. `value` is gs-newton at 10^7."""
    n = args.dense_n
    dev = torch.device("cuda:0")
    tol = float(np.sqrt(np.finfo(np.float32).eps))
    opts = gs.Options(eltype="float32", f_abstol=tol, f_reltol=0.0, max_iterations=200)
    A, _, _ = workloads.tene_batch(n, 1)
    A32 = A.astype(np.float32)
    fam = gs.DenseSin(A32.astype(np.float64))
    prob = gs.NonlinearProblem(fam)
    variants = {"gs-newton": gs.Newton(), "gs-reuse4": gs.Newton(reuse=4),
                "gs-backtracking": gs.Newton(globalization=gs.Backtracking())}
    K = max(args.steps, 3)
    cells, kernel_ms, codes_out, iters_out = {}, {}, {}, {}
    flops = 0.0
    for nsys, label in ((10**6, "1e6"), (10**7, "1e7")):
        rng = np.random.default_rng(1)
        xs = (2.0 * rng.random((nsys, n), dtype=np.float64) - 1.0).astype(np.float32)
        B = (xs @ A32.T + np.sin(xs) / np.float32(10)).astype(np.float32)  # b = A x* + sin.(x*)/10 in Float32
        p = gs.PerSystem(torch.as_tensor(B, device=dev))
        del xs
        S = gs.status_buffer(nsys, dev, "float32")
        for name, m in variants.items():
            gs.solve_batch(torch.zeros((nsys, n), dtype=torch.float32, device=dev), S, prob, m, p, opts)
            torch.cuda.synchronize()
            times = []
            for _ in range(K):
                X = torch.zeros((nsys, n), dtype=torch.float32, device=dev)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                gs.solve_batch(X, S, prob, m, p, opts)
                times.append(time.perf_counter() - t0)
            med = float(np.median(times))
            st = gs.statuses(S)
            cells.setdefault(name, {})[label] = nsys / med
            kernel_ms.setdefault(name, {})[label] = med * 1e3
            codes_out.setdefault(name, {})[label] = np.bincount(st["code"], minlength=6).tolist()
            iters_out.setdefault(name, {})[label] = float(st["iterations"].mean())
            if name == "gs-newton" and label == "1e7":
                flops = dense_flops_per_iter(n) * float(st["iterations"].sum()) / med / 1e12
            del X
        del S, p, B
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        k = 1 << 15
        rng = np.random.default_rng(1)
        xs = (2.0 * rng.random((k, n)) - 1.0).astype(np.float32)
        Bk = (xs @ A32.T + np.sin(xs) / np.float32(10)).astype(np.float32)
        reps, cel, (_, so) = _timed_oracle(
            lambda: orc.solve_batch(gs.Newton(), fam, np.zeros((k, n), dtype=np.float32), gs.PerSystem(Bk), opts,
                                    nthreads=threads), 8.0)
        cpu = {"value": k * reps / cel, "unit": "systems/s", "cores": threads, "kind": "port",
               "sample": f"gs-newton in Float32 on {k} systems of the same kind x {reps} repetitions, {cel:.1f} s "
                         "wall, OpenMP schedule(static) over systems"}
    ref = {m: REFERENCE_GRID[m][n] for m in REFERENCE_GRID if n in REFERENCE_GRID[m]}
    return {
        "metric": "systems/s, the §10.E R2 device grid in Float32 (benchmark/devices.jl), 1 GPU",
        "value": cells["gs-newton"]["1e7"],
        "unit": "systems/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": 1,
        "ms_per_step": kernel_ms["gs-newton"]["1e7"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (A = 2I + R from seed n, x* ~ U[-1,1]^n, numpy PCG64; rounded to Float32)",
        "config": {"workload": f"dense_grid (§10.E A x + sin(x)/10 - b) n={n}, Float32, x0 = 0, abstol sqrt(eps(Float32)), "
                               "reltol 0, max_iterations 200, status form",
                   "cells_systems_per_s": cells, "launch_ms": kernel_ms, "codes": codes_out,
                   "mean_iterations": iters_out,
                   "reference_published": ref,
                   "vs_reference": {m: {v: {c: cells[v][c] / r[v][c] for c in r[v]} for v in r}
                                    for m, r in ref.items()}},
        "roofline": {
            "bound": "fp32 valu",
            "kernel": (f"dense_kernel<float, {n}, Static, status>" if n > 8 else
                       f"solve_small_kernel<DenseSin {n}, Static, Float32, status>") +
                      ", gs-newton at 10^7 systems; achieved = (2/3 n^3 + 4 n^2) flop per Newton iteration x the "
                      "batch's iterations / launch time",
            "achieved": flops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": flops / FP32_PEAK_TFLOPS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }


def sub_results():
    """Compact lines of the other configurations (the n = 48 stage systems, the uneven least-squares
    batch, the dense family at n = 16 and n = 64, the Float32 device grid at n = 16), each measured by its own `--workload` run in a child process (fewer
    steps, no CPU baseline), one after the other; a child that fails ends the series, and its exit status
    is recorded in place of its line."""
    out = {}
    for name, w in (("fpu_gl3", ["fpu_gl3"]), ("expsum_lsq", ["expsum_lsq"]),
                    ("dense_sin_n16", ["dense_sin", "--dense-n", "16"]), ("dense_sin_n64", ["dense_sin", "--dense-n", "64"]),
                    ("dense_grid_n16", ["dense_grid", "--dense-n", "16"])):
        cmd = [sys.executable, str(ROOT / "bench.py"), "--workload", *w, "--steps", "3", "--warmup", "1",
               "--no-cpu-baseline"]
        print(f"bench: sub-result {name}", file=sys.stderr, flush=True)
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        except subprocess.TimeoutExpired:
            out[name] = {"error": "timed out after 300 s"}
            break
        if r.returncode != 0:
            out[name] = {"error": f"exit status {r.returncode}", "stderr_tail": r.stderr[-1500:]}
            break
        line = json.loads(r.stdout.strip().splitlines()[-1])
        out[name] = {"value": line["value"], "unit": line["unit"], "ms_per_step": line["ms_per_step"],
                     "roofline_frac": line["roofline"]["frac"], "roofline_bound": line["roofline"]["bound"],
                     "workload": line["config"]["workload"]}
        if "plain_value" in line["config"]:
            out[name]["plain_value"] = line["config"]["plain_value"]
        if "cells_systems_per_s" in line["config"]:
            out[name]["cells_systems_per_s"] = line["config"]["cells_systems_per_s"]
            out[name]["vs_reference"] = line["config"]["vs_reference"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--workload", choices=("henon_heiles", "fpu_gl3", "expsum_lsq", "dense_sin", "dense_grid"),
                    default="henon_heiles")
    ap.add_argument("--dense-n", type=int, choices=(2, 4, 6, 8, 12, 16, 24, 32, 48, 64), default=16,
                    help="the size of the dense_sin and dense_grid workloads")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--save-net", default=None, help="save the trained network (.npz) of the default workload")
    ap.add_argument("--load-net", default=None, help="load a network saved by --save-net instead of training")
    ap.add_argument("--no-sub-results", action="store_true",
                    help="leave out the compact lines of the other configurations (each measured in a child process)")
    args = ap.parse_args()
    faulthandler.enable()  # a fatal signal prints the Python stack of every thread to stderr
    if args.gpus != 1:
        print("bench.py measures one GPU: the systems are independent and this build runs on a single "
              "MI355X; --gpus above 1 is not supported", file=sys.stderr)
        return 2

    import importlib

    import torch

    import __graft_entry__

    gs = __graft_entry__.load_package()
    workloads = importlib.import_module(gs.__name__ + ".workloads")
    fn = {"henon_heiles": bench_henon_heiles, "fpu_gl3": bench_fpu_gl3, "expsum_lsq": bench_expsum,
          "dense_sin": bench_dense, "dense_grid": bench_dense_grid}[args.workload]
    if args.workload != "henon_heiles" and "--steps" not in sys.argv:
        args.steps, args.warmup = 5, 1
    line = fn(args, gs, workloads, torch)
    if args.workload == "henon_heiles" and not args.no_sub_results:
        line["config"]["sub_results"] = sub_results()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
