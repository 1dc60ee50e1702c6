#!/usr/bin/env python
"""PLL f+grad evaluations per second at L=250, N=5000 on one MI355X.

A step is one call of the evaluate(instance, x, g, nvar) callback on
device-resident x and g (f read back to the host), for the seeded
synthetic MSA with L=250 columns, N=5000 sequences and 60 planted contacts
(seed 2), reweighting threshold 0.8, lambda_single 0.01,
lambda_pair 0.2*(L-1), and x = init_bias plus symmetric N(0, 0.01^2)
couplings.  `value` is the single-pass BF16-MFMA path (bf16f, the
two-pass forward, is timed beside it); the plain FP32 path, the
OpenMP CPU evaluator (the oracle's restatement of evaluate_cpu_omp.c) and
the single-thread one (of evaluate_cpu.c) are timed in the same run on the
same MSA and x.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

L_COLS, N_SEQS, SEED, N_PAIRS = 250, 5000, 2, 60
MFMA_BF16_PEAK_TFLOPS = 2500.0
MFMA_BF16_SPARSE_PEAK_TFLOPS = 5000.0  # 2:4 structured-sparse BF16 MFMA
HBM_PEAK_GBS = 8000.0


COUNTERS_FILE = os.path.join(ROOT, "profiles", "pll_evaluate_L250_N5000_counters.md")


def measured_counters(kernel_prefix):
    """HBM bytes per launch and MFMA busy fraction of a kernel, from the
    committed rocprofv3 counter summary (FETCH_SIZE / WRITE_SIZE scaled by a
    1 GiB copy); None where the summary has no row for the kernel."""
    try:
        lines = open(COUNTERS_FILE).read().splitlines()
    except OSError:
        return None, None
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) >= 5 and kernel_prefix in cells[0].strip("`"):
            busy = float(cells[4]) if cells[4] else None
            return float(cells[3]), busy
    return None, None


PEAK_FILE = os.path.join(ROOT, "profiles", "mfma_peak_microbench.md")


def measured_mfma_peaks():
    """The register-only MFMA throughputs of the committed microbenchmark
    summary (ccmpred_amd/bin/mfma_peak's JSON line), or {}."""
    try:
        for ln in open(PEAK_FILE).read().splitlines():
            ln = ln.strip()
            if ln.startswith('{"cus"'):
                return json.loads(ln)
    except (OSError, ValueError):
        pass
    return {}


def measured_dense_peak(kernel, achieved):
    """Context beside the published peak for a dense contraction kernel: the
    measured register-only rate of its instruction with operands like its
    own (0/1 against small BF16 values), and the fraction of it achieved."""
    pk = measured_mfma_peaks().get("mfma_32x32x16_bf16_contraction_like_tflops")
    if not pk or kernel == "sparse_fwd_softmax":
        return {}
    return {"peak_measured": pk, "frac_measured": round(achieved / pk, 4),
            "peak_measured_source": os.path.relpath(PEAK_FILE, ROOT)}


LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clock x 2.4 GHz = 78.6 TB/s


def fp32_roofline(step_ms, L, N):
    """The plain FP32 path's gather kernels against their LDS bound: each of
    the N*L*21L coupling additions of a contraction reads one FP32 value
    from LDS (4 bytes), 26.25 GB at L=250, N=5000."""
    if not step_ms:
        return None
    lds_bytes = 4.0 * N * L * 21 * L
    out = {"bound": "lds", "peak": round(LDS_PEAK_GBS, 1), "unit": "GB/s", "algorithmic_bytes": lds_bytes}
    for k in ("gather_fwd", "gather_grad"):
        if k in step_ms:
            ach = lds_bytes / (step_ms[k] * 1e-3) / 1e9
            out[k] = {"avg_ms": round(step_ms[k], 4), "achieved": round(ach, 1),
                      "frac": round(ach / LDS_PEAK_GBS, 4)}
    return out


def make_x(msa):
    from ccmpred_amd.pll import init_bias, nsingle_padded
    L = msa.shape[1]
    x = init_bias(msa).astype(np.float64)
    rng = np.random.Generator(np.random.PCG64(7))
    W = rng.normal(0.0, 0.01, (21, L, 21, L))
    W = W + W.transpose(2, 3, 0, 1)
    idx = np.arange(L)
    W[:, idx, :, idx] = 0.0
    X2 = x[nsingle_padded(L):].reshape(21, L, 32, L)
    X2[:, :, :21, :] += W
    return x.astype(np.float32)


def time_gpu(msa, x_host, precision, steps, warmup, warmup_seconds):
    import torch
    from ccmpred_amd.pll import PLL
    with PLL(msa, precision=precision) as p:
        xd = torch.from_numpy(x_host).cuda()
        gd = torch.empty_like(xd)
        torch.cuda.synchronize()
        # W warm-up steps, continued until warmup_seconds have passed: the
        # GPU raises its clocks over the first tens of milliseconds of load
        # (one evaluation took 725 us at the start of a run and 630 us 25
        # evaluations later), and the timed steps measure the steady state
        t_w = time.perf_counter()
        done = 0
        while done < warmup or time.perf_counter() - t_w < warmup_seconds:
            p.evaluate_ptr(xd.data_ptr(), gd.data_ptr())
            done += 1
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = 0.0
        for _ in range(steps):
            f = p.evaluate_ptr(xd.data_ptr(), gd.data_ptr())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # the per-step split from separate calls: the HIP events between the
        # steps cost host time that the throughput above does not carry
        p.set_timing(True)
        for _ in range(max(10, min(steps, 50))):
            p.evaluate_ptr(xd.data_ptr(), gd.data_ptr())
        ncalls, steps_ms = p.timing()
        p.set_timing(False)
    return dict(evals_per_s=steps / (t1 - t0), ms_per_eval=1e3 * (t1 - t0) / steps, f=f,
                timed_calls=ncalls, step_ms=steps_ms, warmup_steps=done)


def time_cg(msa, precision, iterations):
    """The device CG at L=250, N=5000: `iterations` iterations (epsilon 0)
    from init_bias with the reference's other parameters, vectors on the
    device (the line search on the device for the BF16 forms, on the host
    between evaluations for FP32); the second of two runs is timed (the
    first warms the clocks)."""
    import torch
    from ccmpred_amd.pll import PLL, init_bias
    x0 = init_bias(msa)
    with PLL(msa, precision=precision) as p:
        out = None
        for _ in range(2):
            xd = torch.from_numpy(x0).cuda()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = p.conjugrad(xd, max_iterations=iterations, epsilon=0.0)
            torch.cuda.synchronize()
            out = dict(precision=precision, seconds=round(time.perf_counter() - t0, 4), ret=int(r["ret"]),
                       iterations=int(r["iterations"]), evaluations=int(r["evaluations"]),
                       f=float(r["fx"]))
    out["evaluations_per_s"] = round(out["evaluations"] / out["seconds"], 1)
    return out


def physical_cores():
    """Physical cores among the CPUs this process may run on (one per
    thread_siblings group), and the number of those CPUs."""
    cpus = sorted(os.sched_getaffinity(0))
    groups = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/thread_siblings_list") as fh:
                groups.add(fh.read().strip())
        except OSError:
            groups.add(str(c))
    return len(groups), len(cpus)


def time_cpu(msa, x_host, evals):
    from oracle import oracle as ora
    w = ora.calculate_weights(msa, 0.8)
    L = msa.shape[1]
    # threads: OMP_NUM_THREADS, which cpu_legs sets to the physical cores
    ev = ora.OmpEvaluator(msa, w, 0.01, np.float32(0.2 * (L - 1)))
    g = np.zeros_like(x_host)
    ev(x_host, g)  # warm-up (page faults of the per-call buffers)
    ts = []
    for _ in range(evals):
        t0 = time.perf_counter()
        ev(x_host, g)
        ts.append(time.perf_counter() - t0)
    threads = ev.threads
    ev.close()
    med = float(np.median(ts))
    return dict(evals_per_s=1.0 / med, seconds=ts, threads=threads)


def time_cpu_serial(msa, x_host, evals):
    """The oracle's single-thread restatement of evaluate_cpu.c (seconds per call)."""
    from oracle import oracle as ora
    w = ora.calculate_weights(msa, 0.8)
    L = msa.shape[1]
    ts = []
    for _ in range(evals):
        t0 = time.perf_counter()
        ora.evaluate_f32(msa, w, x_host, 0.01, np.float32(0.2 * (L - 1)))
        ts.append(time.perf_counter() - t0)
    return ts


def run_cpu_child(args):
    """The CPU legs, in a process of their own (see cpu_legs): prints one
    JSON object."""
    from ccmpred_amd.msa import synthetic_msa
    msa, _ = synthetic_msa(L_COLS, N_SEQS, SEED, N_PAIRS)
    x = make_x(msa)
    cpu = time_cpu(msa, x, args.cpu_evals)
    serial = time_cpu_serial(msa, x, args.cpu_serial_evals) if args.cpu_serial_evals > 0 else None
    print(json.dumps({"cpu": cpu, "serial": serial}), flush=True)


def cpu_legs(args):
    """Times the OpenMP and single-thread CPU evaluators in a child process
    started with one OpenMP thread per physical core, bound close
    (OMP_PROC_BIND=close, OMP_PLACES=cores).  A child, because an OpenMP
    runtime that reads those variables at start-up binds the process's
    first thread to one place, and this process has loaded torch's."""
    import subprocess
    cores, _ = HOST_CPUS
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", "--cpu-evals", str(args.cpu_evals),
           "--cpu-serial-evals", str(args.cpu_serial_evals)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, check=True).stdout
    res = json.loads(out.strip().splitlines()[-1])
    return res["cpu"], res["serial"]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def batched_workload(nmsa, seed=999):
    from ccmpred_amd.msa import batched_workload as workload
    return workload(nmsa, seed)


def run_batched(args):
    """Batched form: every MSA solved for `iterations` CG iterations (epsilon
    0, no early stop) from init_bias, all MSAs together; value = MSAs/s over
    the solve (set-up and autotuning reported beside it)."""
    import torch
    from ccmpred_amd.batch import Batch
    from ccmpred_amd.pll import PLL, init_bias, contact_map
    msas = batched_workload(args.batch_size)
    t0 = time.perf_counter()
    b = Batch(msas, precision=args.batch_precision)
    t1 = time.perf_counter()
    print(f"batched: {len(msas)} MSAs set up in {t1 - t0:.1f} s", file=sys.stderr, flush=True)
    tuned = b.autotune(reps=3)
    t2 = time.perf_counter()
    print(f"batched: autotuned in {t2 - t1:.1f} s", file=sys.stderr, flush=True)
    for i, m in enumerate(msas):
        b.set_x(i, init_bias(m))
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    rounds = b.solve(max_iterations=args.iterations, epsilon=0.0, rounds_per_graph=args.rounds_per_graph)
    t4 = time.perf_counter()
    print(f"batched: solved in {t4 - t3:.2f} s ({rounds} rounds)", file=sys.stderr, flush=True)
    res = [b.result(i) for i in range(b.n)]
    # contact maps: formed on the device (only L*L floats copied back) and,
    # for comparison, from x read back to the host
    t5 = time.perf_counter()
    maps = [b.contact_map(i) for i in range(b.n)]
    t6 = time.perf_counter()
    maps_host = [contact_map(b.get_x(i), b.ncol[i]) for i in range(b.n)]
    t7 = time.perf_counter()
    maps_same = sum(int(np.array_equal(u, v)) for u, v in zip(maps, maps_host))
    print(f"batched: contact maps {t6 - t5:.2f} s on the device, {t7 - t6:.2f} s via host x", file=sys.stderr,
          flush=True)
    fwd, grad = b.variants()
    dev_gb = b.device_bytes / 1e9
    b.close()
    flops = sum(r["evaluations"] * 4.0 * m.shape[0] * (21 * m.shape[1]) ** 2 for r, m in zip(res, msas))
    line = {
        "metric": f"batched CG: {len(msas)} MSAs x {args.iterations} iterations, MSAs/s",
        "value": round(len(msas) / (t4 - t3), 4),
        "unit": "MSAs/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"{len(msas)} synthetic MSAs, L uniform [50,600], N log-uniform "
                               "[200,20000], seeds 1000+i, x0 = init_bias",
                   "L": [int(m.shape[1]) for m in msas], "N": [int(m.shape[0]) for m in msas],
                   "iterations": args.iterations, "rounds_per_graph": args.rounds_per_graph},
        "solve_s": round(t4 - t3, 3), "setup_s": round(t1 - t0, 3), "autotune_s": round(t2 - t1, 3),
        "tuned_gemm_ms": [round(v, 3) for v in tuned],
        "precision": args.batch_precision,
        "rounds": rounds, "evaluations": int(sum(r["evaluations"] for r in res)),
        "iterations_done": [r["iterations"] for r in res],
        "status": {str(k): int(sum(r["ret"] == k for r in res)) for k in sorted({r["ret"] for r in res})},
        "effective_tflops": round(flops / (t4 - t3) / 1e12, 1),
        "contact_maps_device_s": round(t6 - t5, 3), "contact_maps_host_s": round(t7 - t6, 3),
        "contact_maps_equal": maps_same,
        "device_gb": round(dev_gb, 2),
        "variants_fwd": fwd.tolist(), "variants_grad": grad.tolist(),
    }
    if args.sequential:
        # the same solves one MSA after another through the single-alignment
        # CG (set-up of each instance outside the timed calls)
        total = 0.0
        same = 0
        for i, m in enumerate(msas):
            with PLL(m, precision=args.batch_precision) as p:
                xd = torch.from_numpy(init_bias(m)).cuda()
                torch.cuda.synchronize()
                ts = time.perf_counter()
                r = p.conjugrad(xd, max_iterations=args.iterations, epsilon=0.0)
                total += time.perf_counter() - ts
                same += int(r["fx"] == res[i]["fx"] and r["iterations"] == res[i]["iterations"])
        print(f"batched: sequential solves in {total:.2f} s", file=sys.stderr, flush=True)
        line["sequential_s"] = round(total, 3)
        line["sequential_msas_per_s"] = round(len(msas) / total, 4)
        line["sequential_same_result"] = same
    print(json.dumps(line), flush=True)


HOST_CPUS = (0, 0)


def main():
    global HOST_CPUS
    # counted before anything loads an OpenMP runtime that might bind this
    # thread to fewer CPUs
    HOST_CPUS = physical_cores()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--warmup-seconds", type=float, default=1.0,
                    help="continue the warm-up steps until this much time has passed")
    ap.add_argument("--forms", default="bf16,bf16f,fp32",
                    help="GPU forms to time beside the headline bf16 (comma list of bf16f, fp32)")
    ap.add_argument("--cg-iterations", type=int, default=50,
                    help="device CG iterations timed at L=250, N=5000 for the bf16 form (0: none)")
    ap.add_argument("--cpu-evals", type=int, default=3)
    ap.add_argument("--cpu-serial-evals", type=int, default=1,
                    help="single-thread evaluate_cpu restatement calls to time (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration)")
    ap.add_argument("--batched", action="store_true",
                    help="time the batched form (many MSAs) instead of single evaluations")
    ap.add_argument("--batch-size", type=int, default=64)
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--rounds-per-graph", type=int, default=8)
    ap.add_argument("--batch-precision", default="bf16", choices=["bf16", "bf16f", "bf16x2"],
                    help="BF16 form of the batched solves (bf16: single-pass forward, the headline "
                         "form; bf16f: two-pass forward)")
    ap.add_argument("--sequential", action="store_true",
                    help="with --batched: also time the same solves one MSA at a time")
    args = ap.parse_args()
    if args.cpu_child:
        run_cpu_child(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path "
              "(the evaluate callback works on one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.batched:
        run_batched(args)
        return

    from ccmpred_amd.msa import synthetic_msa
    msa, _ = synthetic_msa(L_COLS, N_SEQS, SEED, N_PAIRS)
    x = make_x(msa)
    if args.copy_mib:
        # a plain device copy of known size: the reference point that scales
        # the FETCH_SIZE / WRITE_SIZE counters of a profiled run to bytes
        import torch
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    forms = set(args.forms.split(","))
    bf16 = time_gpu(msa, x, "bf16", args.steps, args.warmup, args.warmup_seconds)
    bf16f = (time_gpu(msa, x, "bf16f", args.steps, args.warmup, args.warmup_seconds)
             if "bf16f" in forms else None)
    fp32 = (time_gpu(msa, x, "fp32", max(10, args.steps // 4), max(2, args.warmup // 2),
                     args.warmup_seconds) if "fp32" in forms else None)
    cg = time_cg(msa, "bf16", args.cg_iterations) if args.cg_iterations > 0 else None
    cg32 = time_cg(msa, "fp32", args.cg_iterations) if args.cg_iterations > 0 and "fp32" in forms else None
    cpu, serial = (None, None) if args.no_cpu else cpu_legs(args)

    L, N = L_COLS, N_SEQS
    flops = 2.0 * N * (21 * L) ** 2          # algorithmic FLOPs of one contraction
    # the slower of the two contraction kernels (the forward one carries the
    # column softmax as its epilogue when fused)
    fwd_key = next(k for k in ("sparse_fwd_softmax", "gemm_fwd_softmax", "gemm_fwd") if k in bf16["step_ms"])
    gemm = max((fwd_key, "gemm_grad"), key=lambda k: bf16["step_ms"][k])
    gemm_ms = bf16["step_ms"][gemm]
    achieved = flops / (gemm_ms * 1e-3) / 1e12
    # (the gradient at this size runs as the banded launch, 256-row then
    # 192-row tiles; the single-tile kernel's row where the summary has it)
    prefixes = {"sparse_fwd_softmax": ["k_sparse_fwd<"], "gemm_fwd_softmax": ["k_gemm_onehot_dtl<2,"],
                "gemm_fwd": ["k_gemm_onehot_dtl<0,"],
                "gemm_grad": ["k_gemm_grad_banded<", "k_gemm_onehot_dtl<1,"]}[gemm]
    traffic, mfma_busy = None, None
    for prefix in prefixes:
        traffic, mfma_busy = measured_counters(prefix)
        if traffic is not None:
            break
    fwd_roofline = None
    if fwd_key == "sparse_fwd_softmax":
        # the forward on structured-sparse MFMA: the dense algorithmic FLOPs
        # against the dense peak, and the instruction work it executes (7
        # groups of 4 K slots per coupling column, Lp = 256 columns, 5120
        # sequence rows) against the sparse peak
        fms = bf16["step_ms"][fwd_key]
        executed = 2.0 * 5120 * (21 * 256) * (28 * 256)
        f_traffic, f_busy = measured_counters("k_sparse_fwd<")
        fwd_roofline = {"kernel": fwd_key, "avg_ms": round(fms, 4), "algorithmic_flops": flops,
                        "achieved_dense_equivalent": round(flops / (fms * 1e-3) / 1e12, 2),
                        "peak_dense": MFMA_BF16_PEAK_TFLOPS,
                        "frac_dense": round(flops / (fms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4),
                        "executed_flops": executed,
                        "achieved_executed": round(executed / (fms * 1e-3) / 1e12, 2),
                        "peak_sparse": MFMA_BF16_SPARSE_PEAK_TFLOPS,
                        "frac_sparse": round(executed / (fms * 1e-3) / 1e12 / MFMA_BF16_SPARSE_PEAK_TFLOPS, 4),
                        "unit": "TFLOP/s", "traffic": f_traffic, "mfma_busy": f_busy}
        sp = measured_mfma_peaks().get("smfmac_32x32x32_bf16_contraction_like_tflops")
        if sp:
            fwd_roofline["peak_sparse_measured"] = sp
            fwd_roofline["frac_sparse_measured"] = round(executed / (fms * 1e-3) / 1e12 / sp, 4)
    line = {
        "metric": "PLL f+grad evals/s at L=250,N=5000: BF16-MFMA vs FP32 GPU vs host CPU",
        "value": round(bf16["evals_per_s"], 2),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bf16["ms_per_eval"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "PLL evaluate, synthetic MSA L=250 N=5000 (seed 2, 60 planted pairs)",
                   "L": L, "N": N, "theta": 0.8, "lambda_single": 0.01,
                   "lambda_pair": 0.2 * (L - 1), "x": "init_bias + symmetric N(0,0.01^2) couplings"},
        "roofline": {"bound": "mfma", "kernel": gemm, "achieved": round(achieved, 2),
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": traffic,
                     "algorithmic_flops": flops, "avg_ms": round(gemm_ms, 4),
                     "mfma_busy": mfma_busy,
                     "counters": os.path.relpath(COUNTERS_FILE, ROOT) if traffic else None,
                     **measured_dense_peak(gemm, achieved)},
        "roofline_forward": fwd_roofline,
        "warmup_steps_run": bf16["warmup_steps"],
        "bf16_step_ms": {k: round(v, 4) for k, v in bf16["step_ms"].items()},
        "f_bf16": bf16["f"],
    }
    if cg is not None:
        line["cg_bf16"] = cg
    if cg32 is not None:
        line["cg_fp32"] = cg32
    if bf16f is not None:
        line["bf16f_evals_per_s"] = round(bf16f["evals_per_s"], 2)
        line["bf16f_step_ms"] = {k: round(v, 4) for k, v in bf16f["step_ms"].items()}
    if fp32 is not None:
        line["fp32_evals_per_s"] = round(fp32["evals_per_s"], 2)
        line["bf16_over_fp32"] = round(bf16["evals_per_s"] / fp32["evals_per_s"], 3)
        line["fp32_roofline"] = fp32_roofline(fp32["step_ms"], L, N)
        line["fp32_step_ms"] = {k: round(v, 4) for k, v in fp32["step_ms"].items()}
        line["f_fp32"] = fp32["f"]
    if cpu is not None:
        line["cpu_baseline"] = {"value": round(cpu["evals_per_s"], 4), "unit": "evals/s",
                                "cores": cpu["threads"], "kind": "port",
                                "sample": f"{len(cpu['seconds'])} evaluations of the full L=250,N=5000 "
                                          "workload (median), OpenMP restatement of evaluate_cpu_omp.c, "
                                          "one thread per physical core, OMP_PROC_BIND=close"}
        line["bf16_over_cpu"] = round(bf16["evals_per_s"] / cpu["evals_per_s"], 1)
        line["cpu_model"] = cpu_model()
        line["cpu_physical_cores"], line["cpu_logical_available"] = HOST_CPUS
    if serial is not None:
        line["cpu_single_thread"] = {"value": round(1.0 / float(np.median(serial)), 4), "unit": "evals/s",
                                     "cores": 1, "kind": "port",
                                     "sample": f"{len(serial)} evaluation(s) of the full L=250,N=5000 "
                                               "workload, single-thread restatement of evaluate_cpu.c"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
