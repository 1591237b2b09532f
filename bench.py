"""Headline benchmark: batched L^p-annealed projected-Adam search at n = 8400.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A bench "step" is one launch of the fused local-search kernel over the whole
batch: P = 512 structured-prior candidates at n = 8400, 200 Adam iterations
each with p annealed 4 -> 1e4 and lr 0.085 -> 0.0015 (the reference's cold
tier schedule, src/main.cu:627).  value = P * 200 / (seconds per step)
candidate-steps/s.  Inputs are resident in HBM before the timed region.

Beside it: the plain form (one candidate per launch, same kernel, a bounded
sample of candidates run back to back on the stream), the roofline of the
local-search kernel against the 157.3 TFLOP/s dense FP32 matrix peak, and the
CPU baseline: the reference's own fp64 CPU twin of the kernel (tools/probe.cpp
--opt, built from source into oracle/_ref/) timed on this host, or the oracle
port when that binary is absent.  Then F-evaluations per second: the fp64
evaluation kernel over 512 candidates at n = 512, 4096 and 8400, beside the
reference's long-double OpenMP verifier (tools/probe.cpp --verify) on the same
host.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "candidate-steps/s at n=8400, batched vs plain, F within 1e-12 of FP64"
PEAK_FP32 = 157.3  # TFLOP/s, MI355X dense FP32 (matrix and vector)
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8400)
    ap.add_argument("--P", type=int, default=512)
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--seed", type=int, default=20260905)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0,
                    help="workgroup size of the local search at this n (0: the library's rule)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner and launch with the built-in configuration")
    ap.add_argument("--tune-only", default="",
                    help="comma-separated lengths: run the autotuner at each (P candidates "
                         "of that length, 50-step trials) and print its tables")
    ap.add_argument("--fw-only", action="store_true",
                    help="time the Frank-Wolfe / Chambolle-Pock kernel alone (the Frank-Wolfe ladder's "
                         "levels) and print its JSON")
    ap.add_argument("--polish-only", action="store_true",
                    help="time the whole-device fp64 polish alone and print its JSON")
    ap.add_argument("--f-evals-only", action="store_true",
                    help="run only the F-evaluation leg and print its rows as one JSON line")
    ap.add_argument("--phase-clocks", action="store_true",
                    help="load the profiling build of the library (make clocks) and print the "
                         "local search's per-phase shader-clock breakdown at --n/--P/--iters")
    ap.add_argument("--no-moments", action="store_true",
                    help="--phase-clocks without the moments' buffers (the correlation output in "
                         "LDS)")
    ap.add_argument("--profile-copy-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (calibrates the HBM "
                         "byte counters of a profiler run)")
    return ap.parse_args()


def traffic_from_profiles(n: int, P: int, iters: int):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    f = REPO / "profiles" / f"local_search_n{n}_pmc.md"
    if not f.exists():
        return None
    m = re.search(r"TRAFFIC_BYTES_PER_LAUNCH\s+P=(\d+)\s+iters=(\d+)\s+bytes=([0-9.eE+]+)",
                  f.read_text())
    if not m or int(m.group(1)) != P or int(m.group(2)) != iters:
        return None
    return float(m.group(3))


def cpu_baseline(n: int, seconds: float, gpu_pop, gpu_F):
    """CPU leg: the reference's fp64 twin (or the oracle port) on a bounded
    sample of the same workload, plus the fp64 F check of the GPU results."""
    sys.path.insert(0, str(REPO / "oracle"))
    import numpy as np
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    # F check: GPU fp64 F of every candidate vs the oracle (hostF order), on a
    # pool of host threads (the oracle call releases the GIL)
    from concurrent.futures import ThreadPoolExecutor
    rows = [gpu_pop[p].astype(np.float64) for p in range(gpu_pop.shape[0])]
    with ThreadPoolExecutor(threads) as ex:
        want = np.array(list(ex.map(O.F_hostorder, rows)))
    rel = float(np.max(np.abs(gpu_F - want) / want))
    fcheck = {"max_rel_err": rel, "checked": len(rows),
              "bitwise_equal": int(np.sum(gpu_F == want)), "oracle": "hostF order fp64"}

    probe = REPO / "oracle" / "_ref" / "probe"
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    if probe.exists():
        # calibrate: one iteration of every restart, then size the run
        t0 = time.perf_counter()
        subprocess.run([str(probe), "--opt", "--n", str(n), "--restarts", str(threads),
                        "--iters", "2"], capture_output=True, env=env, check=True)
        per = (time.perf_counter() - t0) / 3.0  # 2 iterations + final objective per restart
        iters = max(2, int(seconds / max(per, 1e-6)))
        t0 = time.perf_counter()
        subprocess.run([str(probe), "--opt", "--n", str(n), "--restarts", str(threads),
                        "--iters", str(iters)], capture_output=True, env=env, check=True)
        dt = time.perf_counter() - t0
        steps = threads * iters
        return {"value": steps / dt, "unit": "candidate-steps/s", "cores": threads,
                "kind": "reference",
                "sample": f"tools/probe.cpp --opt (fp64 CPU twin of the kernel), n={n}, "
                          f"{threads} restarts x {iters} iterations, OpenMP {threads} threads, "
                          f"{dt:.1f} s wall"}, fcheck
    W = importlib.import_module("sidon-autocorrelation-bound_amd.workloads")
    pop = W.structured_prior(threads, n, seed=1).astype(np.float64)
    t0 = time.perf_counter()
    O.local_search_batch(pop, 1, 4.0, 4.0, 0.05, 0.05, nthreads=threads)
    per = time.perf_counter() - t0
    iters = max(1, int(seconds / max(per, 1e-6)))
    t0 = time.perf_counter()
    O.local_search_batch(pop, iters, 4.0, 1e4, 0.085, 0.0015, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": threads * iters / dt, "unit": "candidate-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle fp64 local search, n={n}, {threads} candidates x {iters} "
                      f"iterations, {dt:.1f} s"}, fcheck


# the longest fp32 length whose correlation output fits the LDS beside a and
# the weights; longer ones keep it in global memory (the moments' buffer)
LS_ONCHIP_N = 9296


def ls_by_n(S, W, torch, seed: int, P: int = 512, iters: int = 50):
    """The batched local search at n = 512 ... 13216: P structured-prior
    candidates of one length, one launch of `iters` steps of the anneal
    (p 4 -> 1e4) after a warm-up, HIP events around it; candidate-steps/s and
    executed matrix work (the kernel's block-step count x 8192 FLOP) against
    the FP32 peak, in the launch configuration the library picks for n
    (10000 and 13216: the correlation output in global memory, 24- and
    16-float segments)."""
    out = []
    stream = torch.cuda.current_stream()
    sched = (4.0, 1e4, 0.085, 0.0015)
    for n in (512, 1024, 2048, 4096, 8400, 10000, 13216):
        start = torch.as_tensor(W.structured_prior(P, n, seed=seed + 7 * n), device="cuda")
        pop = start.clone()
        mom, vel = torch.empty_like(pop), torch.empty_like(pop)
        S.local_search(pop, 5, *sched, mom=mom, vel=vel)  # warm-up
        pop.copy_(start)
        torch.cuda.synchronize()
        S.ls_block_steps(reset=True)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        S.local_search(pop, iters, *sched, mom=mom, vel=vel)
        ev1.record(stream)
        torch.cuda.synchronize()
        sec = ev0.elapsed_time(ev1) / 1e3
        steps = S.ls_block_steps(reset=True)
        tf = steps * 8192.0 / sec / 1e12
        out.append({"n": n, "P": P, "iters": iters, "ms": sec * 1e3,
                    "candidate_steps_per_s": P * iters / sec, "executed_tflops": tf,
                    "frac_fp32_peak": tf / PEAK_FP32,
                    "tile": dict(zip(("block", "depth", "per_cu"), S.get_ls_tile(n)))})
    return out


def phase_clock_leg(S, W, torch, args):
    """The headline launch (P candidates at n, `iters` steps of the anneal) on
    the profiling build of the library: shader-clock cycles per candidate-step
    of each phase (barrier to barrier, as each workgroup's first wave sees
    them), the idle wave-cycles at the end of the two product phases, and the
    matrix-core cycles the executed block steps need (four MFMAs of 32 cycles
    per block step on one SIMD; a workgroup's 16 waves share four SIMDs).
    The launch configuration is the library's built-in one for n."""
    import numpy as np
    n, P, iters = args.n, args.P, args.iters
    sched = (4.0, 1e4, 0.085, 0.0015)
    start = torch.as_tensor(W.structured_prior(P, n, seed=args.seed), device="cuda")
    pop = start.clone()
    # the moments' buffers, as in the headline launch; --no-moments keeps the
    # correlation output in LDS instead (the form without them)
    mom, vel = (None, None) if args.no_moments else (torch.empty_like(pop), torch.empty_like(pop))
    S.local_search(pop, iters, *sched, mom=mom, vel=vel)  # warm-up
    torch.cuda.synchronize()
    S.ls_phase_clocks(reset=True)
    S.ls_block_steps(reset=True)
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(args.steps):
        pop.copy_(start)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        S.local_search(pop, iters, *sched, mom=mom, vel=vel)
        ev1.record(stream)
        torch.cuda.synchronize()
        times.append(ev0.elapsed_time(ev1))
    clk = S.ls_phase_clocks(reset=True)
    blocks = S.ls_block_steps(reset=True)
    steps = clk[6]
    waves = S.get_ls_tile(n)[0] // 64
    per = lambda x: x / steps
    phases = {"autoconvolution": per(clk[0]), "weights": per(clk[1]), "correlation": per(clk[2]),
              "update": per(clk[3])}
    total = sum(phases.values())
    matrix = blocks * 128.0 / 4.0 / steps            # cycles of matrix work per candidate-step
    return {"n": n, "P": P, "iters": iters, "launches": args.steps,
            "correlation_output": "LDS" if args.no_moments else "moments' buffer rows",
            "ms_per_launch": float(np.median(times)),
            "tile": dict(zip(("block", "depth", "per_cu"), S.get_ls_tile(n))),
            "candidate_steps": steps, "candidates": clk[7],
            "cycles_per_candidate_step": phases, "total": total,
            "idle_wave_cycles_per_candidate_step": {"autoconvolution_end": per(clk[4]),
                                                    "correlation_end": per(clk[5]),
                                                    "waves": waves},
            "matrix_cycles_per_candidate_step": matrix,
            "block_steps_per_candidate_step": blocks / steps,
            "matrix_share_of_step": matrix / total}


def f_evals(S, W, torch, with_cpu: bool, seed: int):
    """fp64 F-evaluations per second at n = 512, 4096, 8400: sidon_eval_f64 over
    a batch of 512 * (8400 // n)^2 structured-prior vectors (HIP events), beside the
    reference's long-double verifier (tools/probe.cpp --verify, OpenMP, one
    process per vector as the reference runs it) on this host."""
    import numpy as np
    out = []
    probe = REPO / "oracle" / "_ref" / "probe"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    for n in (512, 4096, 8400):
        # batch of about the same work at every n (512 distinct vectors, tiled)
        P = 512 * max(1, (8400 // n) ** 2)
        pop = torch.as_tensor(W.structured_prior(512, n, seed=seed + n), device="cuda")
        pop = pop.repeat(P // 512, 1)
        stream = torch.cuda.current_stream()
        S.eval_f64(pop)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        ev0.record(stream)
        for _ in range(reps):
            S.eval_f64(pop)
        ev1.record(stream)
        torch.cuda.synchronize()
        gpu = P * reps / (ev0.elapsed_time(ev1) / 1e3)
        # k_eval: n^2 fp64 FMAs (2 n^2 FLOP) per candidate on the vector ALUs
        tf = 2.0 * n * n * gpu / 1e12
        row = {"n": n, "gpu_evals_per_s": gpu, "gpu_batch": P, "gpu_fp64_tflops": tf,
               "frac_fp64_peak": tf / PEAK_FP64}
        if with_cpu and probe.exists():
            sys.path.insert(0, str(REPO / "oracle"))
            import oracle as O
            import tempfile
            a = pop[0].double().cpu().numpy()
            with tempfile.TemporaryDirectory() as td:
                path = os.path.join(td, f"vec{n}.txt")
                O.write_vec(path, a, O.F_hostorder(a))
                env = dict(os.environ, OMP_NUM_THREADS=str(threads))
                k, t0 = 0, time.perf_counter()
                while k < 50 and (k < 3 or time.perf_counter() - t0 < 2.0):
                    subprocess.run([str(probe), "--verify", path], capture_output=True, env=env,
                                   check=True)
                    k += 1
                cpu = k / (time.perf_counter() - t0)
            row.update({"cpu_evals_per_s": cpu, "cpu": f"tools/probe.cpp --verify (long double, "
                        f"OpenMP {threads} threads), {k} runs", "gpu_over_cpu": gpu / cpu})
        if with_cpu:
            hf = hostf_evals(pop[:512].double().cpu().numpy(), threads)
            row.update({"cpu_hostf_evals_per_s": hf, "cpu_hostf": f"hostF-order fp64 "
                        f"(src/main.cu:94-105, oracle restatement), {threads} threads over "
                        f"vectors", "gpu_over_cpu_hostf": gpu / hf})
        out.append(row)
    return out


def fw_legs(S, W, torch, seed: int, rounds: int = 1, cp_iters: int = 200, P: int = 256):
    """Frank-Wolfe rounds solved by Chambolle-Pock over P structured-prior
    candidates at the Frank-Wolfe ladder's levels n = 1050, 2100, 4200, 8400 (a bounded
    sample: `rounds` rounds of `cp_iters` iterations, not the 50 x 2000 of a
    full level), HIP events around the launch.  Matrix work the kernel ran is
    its block-step count (4 x 2048 FLOP each); dense-equivalent work is 4 n^2
    per CP iteration plus 2 n^2 (nT + 1) per round."""
    out = []
    stream = torch.cuda.current_stream()
    for n in (1050, 2100, 4200, 8400):
        start = torch.as_tensor(W.structured_prior(P, n, seed=seed + n), device="cuda")
        pop = start.clone()
        S.frank_wolfe(pop, 1, 10)                      # warm-up
        pop.copy_(start)
        torch.cuda.synchronize()
        S.fw_block_steps(reset=True)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        S.frank_wolfe(pop, rounds, cp_iters)
        ev1.record(stream)
        torch.cuda.synchronize()
        s = ev0.elapsed_time(ev1) / 1e3
        steps = S.fw_block_steps(reset=True)
        dense = P * rounds * (4.0 * n * n * cp_iters + 2.0 * n * n * 13)
        out.append({"n": n, "P": P, "rounds": rounds, "cp_iters": cp_iters, "ms": s * 1e3,
                    "cp_iterations_per_s": P * rounds * cp_iters / s,
                    "executed_tflops": steps * 8192.0 / s / 1e12,
                    "frac_fp32_peak": steps * 8192.0 / s / 1e12 / PEAK_FP32,
                    "dense_equivalent_tflops": dense / s / 1e12,
                    "block_size": 1024 if n >= 2048 else (512 if n >= 512 else 256)})
    return out


def grad_leg(S, W, torch, seed: int, n: int = 8400):
    """The backward kernel of J_p (sidon_lp_grad: b recomputed, the correlation
    adjoint, one workgroup a row) on B structured-prior rows at n, p = 8, 64
    and 1024: median of 5 calls after a warm-up, HIP events on the stream.
    Executed matrix work from the kernel's block-step count (8192 FLOP each;
    correlation steps over zero weights, most of them at p = 1024, are
    skipped), and the dense 3 n^2 FLOP a row beside it; B = 64 (the adjoint
    check's batch) occupies 64 of the 256 CUs, B = 256 one row a CU."""
    import numpy as np
    out = []
    stream = torch.cuda.current_stream()
    for B in (64, 256):
        a = torch.as_tensor(W.structured_prior(B, n, seed=seed + B), device="cuda")
        J = torch.empty(B, dtype=torch.float32, device="cuda")
        g = torch.empty_like(a)
        for p in (8.0, 64.0, 1024.0):
            call = lambda: S._check(S.lib().sidon_lp_grad(a.data_ptr(), n, B, p, J.data_ptr(),
                                                          g.data_ptr(), stream.cuda_stream),
                                    "sidon_lp_grad")
            call()
            torch.cuda.synchronize()
            S.ls_block_steps(reset=True)
            ts = []
            for _ in range(5):
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ev0.record(stream)
                call()
                ev1.record(stream)
                torch.cuda.synchronize()
                ts.append(ev0.elapsed_time(ev1) / 1e3)
            s = float(np.median(ts))
            executed = S.ls_block_steps(reset=True) / 5 * 8192.0 / s / 1e12
            out.append({"n": n, "B": B, "p": p, "ms": s * 1e3, "gradients_per_s": B / s,
                        "executed_tflops": executed, "frac_fp32_peak": executed / PEAK_FP32,
                        "dense_equivalent_tflops": 3.0 * n * n * B / s / 1e12})
    return out


def polish_leg(S, W, torch, seed: int, n: int = 8400, K: int = 8, iters: int = 100):
    """The whole-device fp64 polish (sidon_polish_f64) of K candidates at n over
    `iters` steps of the anneal, HIP events around the launch sequence.  Dense
    matrix work: n^2 / 2 fp64 multiply-adds for the symmetric autoconvolution
    plus n^2 for the correlation, per candidate-step."""
    start = torch.as_tensor(W.structured_prior(K, n, seed=seed).astype("float64"), device="cuda")
    ws = torch.empty(S.lib().sidon_polish_f64_workspace(n, K), dtype=torch.uint8, device="cuda")
    sched = (200.0, 6000.0, 0.01, 0.0002)
    pop = start.clone()
    S.polish_f64(pop, 4, *sched, workspace=ws)            # warm-up
    pop.copy_(start)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    S.polish_f64(pop, iters, *sched, workspace=ws)
    ev1.record(stream)
    torch.cuda.synchronize()
    s = ev0.elapsed_time(ev1) / 1e3
    dense = 2.0 * 1.5 * n * n * K * iters
    return {"n": n, "K": K, "iters": iters, "ms": s * 1e3,
            "candidate_steps_per_s": K * iters / s, "dense_fp64_tflops": dense / s / 1e12,
            "frac_fp64_peak": dense / s / 1e12 / PEAK_FP64,
            "schedule": "p 200 -> 6000, lr 0.01 -> 0.0002 (the driver's first polish pass)"}


def hostf_evals(vecs, threads: int, seconds: float = 1.5) -> float:
    """hostF-order fp64 F-evaluations per second on `threads` host threads, one
    vector per task (the oracle's C restatement; ctypes releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O
    O.F_hostorder(vecs[0])  # load the library before timing
    done, t0 = 0, time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while done < 4 * threads or time.perf_counter() - t0 < seconds:
            batch = [vecs[(done + j) % len(vecs)] for j in range(threads)]
            list(ex.map(O.F_hostorder, batch))
            done += threads
    return done / (time.perf_counter() - t0)


def tune_table(S, args):
    """The autotuner's trials at each length: P structured-prior candidates
    of that length, 50 projected-Adam steps per trial, three timed trials
    after a warm-up; the built-in rule and the tuner's choice beside them."""
    out = {}
    for n in (int(x) for x in args.tune_only.split(",")):
        builtin = S.get_ls_tile(n)
        tile, ms, table = S.tune_local_search(n, P=args.P, iters=50, repeats=3,
                                              seed=args.seed + 1)
        S.set_ls_tile(n, 0)
        name = "{}/{}/{}".format
        out[n] = {"builtin": name(*builtin), "choice": name(*tile), "ms": ms,
                  "builtin_ms": table.get(builtin),
                  "trials_ms": {name(*k): v for k, v in sorted(table.items())}}
        print(f"[tune] n={n}: {out[n]['choice']} {ms:.3f} ms", file=sys.stderr, flush=True)
    return out


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); every box here "
              "has one GPU and this build has no multi-GPU path", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    S = importlib.import_module("sidon-autocorrelation-bound_amd")
    W = importlib.import_module("sidon-autocorrelation-bound_amd.workloads")
    if args.phase_clocks:
        S.use_phase_clock_library()
    S.init(0)
    host = {"nproc": os.cpu_count(),
            "threads": int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)}
    if args.profile_copy_mib and (args.fw_only or args.polish_only or args.f_evals_only):
        src = torch.ones(args.profile_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.tune_only:
        print(json.dumps({"autotune": tune_table(S, args), "host": host}))
        return 0
    if args.phase_clocks:
        print(json.dumps({"phase_clocks": phase_clock_leg(S, W, torch, args), "host": host}))
        return 0
    if args.fw_only:
        print(json.dumps({"frank_wolfe": fw_legs(S, W, torch, args.seed), "host": host}))
        return 0
    if args.polish_only:
        print(json.dumps({"polish": polish_leg(S, W, torch, args.seed), "host": host}))
        return 0
    if args.f_evals_only:
        print(json.dumps({"f_evals": f_evals(S, W, torch, not args.no_cpu, args.seed),
                          "host": host}))
        return 0
    n, P, iters = args.n, args.P, args.iters
    sched = (4.0, 1e4, 0.085, 0.0015)
    tuning = None
    if args.block:
        S.set_ls_block(n, args.block)
    elif not args.no_tune:
        # the autotuner picks waves per candidate, operand depth and candidates
        # a CU at n on the workload's own population size (20-step trials)
        tile, tms, table = S.tune_local_search(n, P=P, iters=20, repeats=2, seed=args.seed + 1)
        tuning = {"choice": {"block": tile[0], "depth": tile[1], "per_cu": tile[2]},
                  "ms_per_trial": tms, "trial_iters": 20,
                  "trials": [{"block": b, "depth": d, "per_cu": c, "ms": m}
                             for (b, d, c), m in sorted(table.items())]}

    if args.profile_copy_mib:
        src = torch.ones(args.profile_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    start = torch.as_tensor(W.structured_prior(P, n, seed=args.seed), device="cuda")
    pop = start.clone()
    stream = torch.cuda.current_stream()
    # the moments' buffers, as the reference's runLocalSearch passes them
    # (src/main.cu:152-159); the 1024-thread launch keeps its correlation
    # output in the rows of the first until the final moments are written
    mom, vel = torch.empty_like(pop), torch.empty_like(pop)

    for _ in range(args.warmup):
        pop.copy_(start)
        S.local_search(pop, iters, *sched, mom=mom, vel=vel)
    torch.cuda.synchronize()

    times = []
    S.ls_block_steps(reset=True)
    for _ in range(args.steps):
        pop.copy_(start)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        S.local_search(pop, iters, *sched, mom=mom, vel=vel)
        ev1.record(stream)
        torch.cuda.synchronize()
        times.append(ev0.elapsed_time(ev1) / 1e3)
    ms = float(np.median(times)) * 1e3
    value = P * iters / (ms / 1e3)
    block_steps = S.ls_block_steps(reset=True) / args.steps  # matrix block steps per launch

    # plain form: one candidate per launch, back to back on the same stream
    k = min(args.plain_sample, P)
    plain_value, same = None, None
    if k > 0:
        plain = start[:k].clone()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        pmom, pvel = torch.empty_like(plain), torch.empty_like(plain)
        for p in range(k):
            S.local_search(plain[p:p + 1], iters, *sched, mom=pmom[p:p + 1],
                           vel=pvel[p:p + 1])
        ev1.record(stream)
        torch.cuda.synchronize()
        plain_value = k * iters / (ev0.elapsed_time(ev1) / 1e3)
        same = bool(torch.equal(plain, pop[:k]))

    # whole-device single candidate: the fp64 search with both products cut
    # into work units for every CU (sidon_polish_f64), same start and schedule
    single = None
    if k > 0:
        one = start[:1].double().contiguous()
        ws = torch.empty(S.lib().sidon_polish_f64_workspace(n, 1), dtype=torch.uint8,
                         device="cuda")
        S.polish_f64(one.clone(), 2, *sched, workspace=ws)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        S.polish_f64(one, iters, *sched, workspace=ws)
        ev1.record(stream)
        torch.cuda.synchronize()
        single = {"value": iters / (ev0.elapsed_time(ev1) / 1e3), "unit": "candidate-steps/s",
                  "form": "one candidate over the whole device in fp64 (sidon_polish_f64: both "
                          "products as fp64 matrix-core work units over every CU, five "
                          "kernels a step)",
                  "dtype": "f64"}

    valu_ceiling, mfma_ceiling = S.measure_fp32_ceilings()

    F, _ = S.eval_f64(pop)
    torch.cuda.synchronize()
    F0, _ = S.eval_f64(start)
    # Matrix work the launch executed, counted by the kernel: block steps of
    # four 16x16x4 MFMAs (8192 FLOP).  The dense symmetric form needs 3 n^2
    # FLOP per candidate-step (each product a_i a_j, i < j, once for b; n^2
    # multiply-adds for G); correlation steps whose weights are all zero are
    # skipped.  The reference counts 4 n^2 (src/main.cu:313, :625).
    exec_flop = block_steps * 8192.0
    achieved = exec_flop / (ms / 1e3) / 1e12
    dense_equiv = 3.0 * n * n * iters * P / (ms / 1e3) / 1e12
    ref_accounting = 4.0 * n * n * iters * P / (ms / 1e3) / 1e12

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "candidate-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"batched anneal: {P} structured-prior candidates at n={n}, "
                               f"{iters} projected-Adam steps per launch, p 4->1e4, "
                               f"lr 0.085->0.0015, one persistent launch",
                   "P": P, "n": n, "iters_per_step": iters, "seed": args.seed,
                   "block": S.get_ls_block(n),
                   "tile": dict(zip(("block", "depth", "per_cu"), S.get_ls_tile(n)))},
        "autotune": tuning,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": traffic_from_profiles(n, P, iters),
                     "executed_flop_per_launch": exec_flop,
                     "dense_flop_per_candidate_step": 3.0 * n * n,
                     "dense_equivalent_tflops": dense_equiv,
                     "reference_accounting_tflops": ref_accounting},
        "plain": {"value": plain_value, "unit": "candidate-steps/s", "sample": k,
                  "form": "one candidate per launch (grid = 1 workgroup), launched back to "
                          "back on one stream", "bitwise_equal_to_batched": same},
        "batched_over_plain": value / plain_value if plain_value else None,
        "single_candidate_whole_device": single,
        "engine_ceilings": {"valu_fp32_tflops": valu_ceiling, "mfma_fp32_tflops": mfma_ceiling,
                            "note": "register-only FMA kernels on every CU, this build's flags "
                                    "(packed FP32 off); the local search runs on the matrix "
                                    "cores"},
        "F_mean_before": float(F0.mean().item()),
        "F_mean_after": float(F.mean().item()),
    }
    if not args.no_cpu:
        cpu, fcheck = cpu_baseline(n, args.cpu_seconds, pop.cpu().numpy(), F.cpu().numpy())
        out["cpu_baseline"] = cpu
        out["f_check"] = fcheck
    out["ls_by_n"] = ls_by_n(S, W, torch, args.seed)
    out["f_evals"] = f_evals(S, W, torch, not args.no_cpu, args.seed)
    out["lp_grad"] = grad_leg(S, W, torch, args.seed)
    out["host"] = host
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
