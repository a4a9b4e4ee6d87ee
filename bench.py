"""Benchmark: time to 16 roots at 1e-8 Eh with FP32 sigma (switching to FP64)
against the all-FP64 GPU run, and the sigma kernel's HBM roofline fraction.

One "step" is one complete mixed-precision block-Davidson solve for the
lowest 16 roots, with the space, integrals and (for the stored workload) the
Hamiltonian already resident in HBM; on the direct workload the subspace
holds up to 8 x 16 = 128 vectors.  The all-FP64 solve of the same space,
with the same solver settings, is run once beside it; every root of each
timed solve must agree with it to 1e-8 Eh or the run fails.

Prints ONE JSON line (rank 0, N = 1).  Usage:
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mrci10x24]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

T_START = time.time()  # the wall budget counts from process start (torch import included)
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}  # MI355X dense MFMA
METRIC = "time to 16 roots at 1e-8 Eh, FP32-sigma vs FP64 on GPU; sigma HBM GB/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed mixed solves, at most (default 3 direct, 5 stored); capped by --budget")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed mixed solves, at most (default 0 direct, 1 stored); capped by --budget")
    ap.add_argument("--budget", type=float, default=420.0,
                    help="wall seconds from process start within which every solve must end; a solve starts "
                         "only if the longest one so far still fits, with the tail (sigma timings, CPU leg) "
                         "reserved; at least one timed solve always runs")
    ap.add_argument("--workload", default="mrci10x24")
    ap.add_argument("--engine", type=int, default=0,
                    help="direct sigma engine: 0 auto (string-driven where the space is complete RAS), "
                         "1 configuration pair records, 2 string-driven")
    ap.add_argument("--roots", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--guard-roots", type=int, default=0,
                    help="guard roots in the FP32 phase of the mixed solve")
    ap.add_argument("--max-subspace", type=int, default=None,
                    help="Davidson subspace cap in vectors (default 8 x roots direct, the library's 5 x roots "
                         "stored); the same for the FP64 and the mixed solves")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--graph", type=int, default=0,
                    help="Davidson steps: 0 stored-sigma steps as captured hipGraphs, 1 eager launches")
    ap.add_argument("--select-rule", type=int, default=0,
                    help="Davidson corrections per iteration: 0 the padded kernel width's free columns take more, "
                         "1 the reference's one per unconverged root")
    ap.add_argument("--precond", type=int, default=0,
                    help="Davidson preconditioner: 0 the reference's diagonal one, 1 the same with Olsen's "
                         "correction")
    ap.add_argument("--sigma-reps", type=int, default=50)
    ap.add_argument("--sigma-k", type=int, default=None,
                    help="--sigma-only: block width to time (default: the roots)")
    ap.add_argument("--sigma-only", action="store_true",
                    help="time sigma launches only (no Davidson); for profiling runs")
    ap.add_argument("--subspace-products", action="store_true",
                    help="time the Davidson subspace products (psoci_project, psoci_combine) over the workload's "
                         "SEF count at basis widths 16..128 and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain 1 GiB device copy (reference reading for the FETCH/WRITE counters)")
    return ap.parse_args()


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def time_sigma(H, api, torch, k, prec, reps, mode, warm=3):
    """Average duration of one sigma launch set (all kernels of one call),
    timed with HIP events recorded on the library's own stream."""
    n = H.n
    dt = torch.float32 if prec == api.FP32 else torch.float64
    X = torch.randn(n, k, device="cuda", dtype=torch.float64).to(dt).contiguous()
    Y = torch.empty_like(X)
    stream = torch.cuda.ExternalStream(H.stream_handle())
    for _ in range(warm):
        H.sigma_device(X, Y, mode, prec)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        H.sigma_device(X, Y, mode, prec)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps  # ms


def measured_traffic(workload, k):
    """HBM bytes per launch of the dominant kernel on this workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/ (each
    scaled by the same counter's reading for a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    rec = table.get(f"{workload}/k{k}")
    if not rec:
        return None
    return {"bytes_per_launch": int(rec["bytes_per_launch"]), "source": rec["source"]}


def cpu_stored(args, cfg, ints, k, ref):
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    sp = oracle.OracleSpace(cfg, ints)
    tc = time.time()
    sp.build_h(threads)
    tb = time.time() - tc
    tc = time.time()
    # the same starting vectors as the GPU solves (unit vectors on the k lowest
    # diagonal elements, guess 0)
    oo = sp.davidson(k, rtol=args.tol, maxit=4000, guess=0)
    td = time.time() - tc
    diff = float(np.abs(oo["evals"] - ref["evals"]).max())
    valid = diff <= 1e-8
    return {"value": round(tb + td, 3), "unit": "s", "cores": threads, "kind": "port",
            "sample": (f"whole solve, not sampled: oracle (C++ restatement of the reference CPU path) on "
                       f"{args.workload}: H build on {threads} OpenMP threads {tb:.2f} s + reference-schedule "
                       f"single-vector Davidson (1 thread) for {k} roots to residual {args.tol:g} "
                       f"({oo['iterations']} iterations) {td:.2f} s; host {cpu_model()}"
                       + ("" if valid else
                          "; INVALID as a baseline: the single-vector schedule converged to other roots than the "
                          "GPU solve (it corrects one root per iteration from unit guesses and can miss a copy of "
                          "a degenerate eigenvalue)")),
            "max_root_diff_vs_gpu_fp64": diff, "valid": valid}


def hbuild_measure(api, torch, with_cpu, reps=3):
    """The stored H build (psoci_build_h: pair screening, element blocks,
    thresholded packing into CSR) of the 20-spin-orbital / 8-electron full CI
    space (fci10, 31,650 SEFs): GPU time per build after one warm-up build,
    elements (lower-triangle nonzeros incl. the diagonal) per second and
    output bytes nnz_L * 12 + (n + 1) * 8 per second; beside it the oracle's
    OpenMP build of the same matrix on the host."""
    from psoci_amd import inputs
    cfg, ints, _ = inputs.workload("fci10")
    H = api.Hamiltonian(cfg, ints)
    nnz = H.build()
    torch.cuda.synchronize()
    ts, tt = [], []
    for _ in range(reps):
        nnz = H.build()
        bt = H.build_timing()
        ts.append(bt["build_s"])
        tt.append(bt["tune_s"])
    t = float(np.median(ts))
    n = H.n
    out_bytes = nnz * 12 + (n + 1) * 8
    rec = {"workload": "fci10 (20 spin-orbitals, 8 electrons, 31,650 SEFs)", "nnz_lower": int(nnz),
           "gpu_ms": round(t * 1e3, 3), "elements_per_s": round(nnz / t, 1),
           "output_bytes": int(out_bytes), "output_gbs": round(out_bytes / t / 1e9, 2),
           "tune_ms": round(float(np.median(tt)) * 1e3, 3),
           "note": ("median of 3 builds after a warm-up build (psoci_build_h_timing): every build screens the "
                    "pairs, generates the blocks and packs the CSR anew; the stored sigma's shape autotuner that "
                    "follows each build is tune_ms")}
    if with_cpu:
        import oracle
        threads = min(os.cpu_count() or 1, 16)
        sp = oracle.OracleSpace(cfg, ints)
        t0 = time.time()
        nnz_o = sp.build_h(threads)
        tc = time.time() - t0
        rec["cpu"] = {"seconds": round(tc, 3), "threads": threads, "kind": "port",
                      "nnz_lower": int(nnz_o), "nnz_matches": int(nnz_o) == int(nnz),
                      "gpu_speedup": round(tc / t, 1)}
    return rec


def cpu_direct(args, cfg, ints, k, sigma_calls, budget_s=15.0):
    """Sampled rows of the reference direct sigma (every block regenerated),
    timed until about budget_s seconds of wall time, scaled to the solve."""
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    sp = oracle.OracleSpace(cfg, ints)
    rng = np.random.default_rng(11)
    X = rng.standard_normal((sp.n, k))
    done, tot = 0, 0.0
    while tot < budget_s and done < cfg.ncfg:
        rows = rng.choice(cfg.ncfg, threads * 2, replace=False)
        tc = time.time()
        sp.direct_rows(rows, X, threads)
        tot += time.time() - tc
        done += len(rows)
    per_sigma = tot / done * cfg.ncfg
    return {"value": round(per_sigma * sigma_calls, 1), "unit": "s", "cores": threads, "kind": "port",
            "extrapolated": True,
            "sample": (f"oracle direct sigma (every block regenerated, the reference's direct path) on {done} "
                       f"random configuration rows of {args.workload}, {k} vectors, {threads} OpenMP threads: "
                       f"{tot:.1f} s; scaled by ncfg/{done} to {per_sigma:.0f} s per sigma and by the "
                       f"{sigma_calls} sigma calls of the GPU mixed solve (Davidson vector work not counted); "
                       f"host {cpu_model()}")}


def subspace_products(api, inputs, torch, n, workload):
    """Wall time of the Davidson subspace products over n rows of a basis of
    128 FP64 columns (row stride 128, as the solve's V and HV): the
    projection A^T B (psoci_project, B 16 columns) and the update A Y
    (psoci_combine, 16 output columns) at basis widths m = 16..128, best of 3
    synchronised calls; bytes = n (m + 16) 8."""
    cfg6, ints6, _ = inputs.workload("fci6")
    H = api.Hamiltonian(cfg6, ints6)  # a context; the products take any n
    V = torch.randn(n, 128, dtype=torch.float64, device="cuda")
    W = torch.randn(n, 128, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rng = np.random.default_rng(7)
    out = {"workload": workload, "n": n, "ld": 128, "project": [], "combine": []}
    for m in (16, 32, 48, 64, 80, 96, 112, 128):
        best = min(_timed(lambda: H.project(V[:, :m], W[:, 64:80])) for _ in range(3))
        out["project"].append({"m": m, "b": 16, "ms": round(best * 1e3, 4),
                               "gbs": round(n * (m + 16) * 8 / best / 1e9, 1)})
        Y = rng.random((m, 16))
        best = min(_timed(lambda: H.combine(V[:, :m], Y, W[:, 0:16])) for _ in range(3))
        out["combine"].append({"m": m, "k": 16, "ms": round(best * 1e3, 4),
                               "gbs": round(n * (m + 16) * 8 / best / 1e9, 1)})
    print(json.dumps(out))


def _timed(fn):
    t = time.perf_counter()
    fn()
    return time.perf_counter() - t


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from psoci_amd import api, inputs
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(3)
    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 3}))
        return

    cfg, ints, _ = inputs.workload(args.workload)
    direct = inputs.WORKLOADS[args.workload]["kind"] != "fci"
    if args.subspace_products:
        subspace_products(api, inputs, torch, int(cfg.nsef().sum()), args.workload)
        return
    if args.steps is None:
        args.steps = 3 if direct else 5
    if args.warmup is None:
        args.warmup = 0 if direct else 1
    args.steps = max(1, args.steps)
    args.warmup = max(0, args.warmup)
    mode = api.DIRECT if direct else api.STORED
    k = args.roots
    guard = args.guard_roots
    msub = args.max_subspace if args.max_subspace is not None else (8 * k if direct else 0)
    n_sef = int(cfg.nsef().sum())
    print(f"[bench] {args.workload}: ncfg={cfg.ncfg} nsef={n_sef} mode={'direct' if direct else 'stored'} k={k}",
          file=sys.stderr, flush=True)
    # the inputs travel as the reference's do: the configuration list and a
    # SIFS-style integral file, written here and read back by the library
    # (outside the timed region)
    import tempfile
    in_dir = tempfile.mkdtemp(prefix="psoci_bench_")
    sifs_path = os.path.join(in_dir, f"{args.workload}.sifs")
    cfg_path = os.path.join(in_dir, f"{args.workload}.configs")
    inputs.write_sifs(sifs_path, ints)
    inputs.write_configs(cfg_path, cfg)
    t0 = time.time()
    H = api.Hamiltonian.from_files(cfg_path, sifs_path)
    if not (np.array_equal(H.cfg.occ, cfg.occ) and np.array_equal(H.cfg.jsym, cfg.jsym)):
        raise SystemExit("configuration file read back differs from the list written")
    os.remove(sifs_path)
    os.remove(cfg_path)
    os.rmdir(in_dir)
    engine, ras = None, None
    if direct:
        engine = H.direct_engine(args.engine)
        if engine == 1:
            npairs, nelems = H.direct_info()
        else:
            ras = H.direct_ras_info()
        nnz = None
    else:
        nnz = H.build()
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"[bench] setup {setup_s:.2f} s", file=sys.stderr, flush=True)
    if args.sigma_only:
        reps = 2 if direct else args.sigma_reps
        ks = args.sigma_k or k
        ms32 = time_sigma(H, api, torch, ks, api.FP32, reps, mode, 1)
        ms64 = time_sigma(H, api, torch, ks, api.FP64, reps, mode, 1)
        print(json.dumps({"workload": args.workload, "k": ks, "sigma_fp32_ms": round(ms32, 4),
                          "sigma_fp64_ms": round(ms64, 4), "setup_seconds": round(setup_s, 3)}))
        return

    # all-FP64 reference solve of the same space
    ref = H.davidson(k, prec=api.FP64, mode=mode, tol=args.tol, want_vectors=False, precond=args.precond,
                     graph=args.graph, select_rule=args.select_rule, max_subspace=msub)
    if not ref["converged"]:
        raise SystemExit("FP64 Davidson did not converge")
    fp64_s = ref["stats"]["seconds"]
    print(f"[bench] fp64 solve {fp64_s:.2f} s, {ref['stats']['iterations']} iterations", file=sys.stderr, flush=True)

    # wall budget: what follows the solves (sigma timings, CPU leg) is
    # reserved; a solve starts only if the longest solve so far still fits.
    # Warm-up solves leave room for min(steps, 2) timed ones.
    tail_s = 60.0 if direct else 30.0
    if args.no_cpu:
        tail_s = 20.0 if direct else 5.0
    durations = []  # every mixed solve run so far; before the first, the FP64 solve's time is the estimate

    def fits(n_more):
        est = max(durations) if durations else fp64_s
        return time.time() - T_START + n_more * est + tail_s <= args.budget

    warm_done = 0
    while warm_done < args.warmup and fits(1 + min(args.steps, 2)):
        tw = time.time()
        H.davidson(k, prec=api.MIXED, mode=mode, tol=args.tol, switch_tol=1e-4, want_vectors=False,
                   extra_roots=guard, precond=args.precond, graph=args.graph, select_rule=args.select_rule,
                   max_subspace=msub)
        torch.cuda.synchronize()
        durations.append(time.time() - tw)
        warm_done += 1
        print(f"[bench] warm-up solve {warm_done}: {durations[-1]:.2f} s", file=sys.stderr, flush=True)
    times, dmax, last = [], 0.0, None
    t_all = time.time()
    for s in range(args.steps):
        if s > 0 and not fits(1):
            print(f"[bench] wall budget {args.budget:.0f} s reached after {s} timed solves", file=sys.stderr,
                  flush=True)
            break
        torch.cuda.synchronize()
        ts = time.time()
        out = H.davidson(k, prec=api.MIXED, mode=mode, tol=args.tol, switch_tol=1e-4, want_vectors=False,
                         extra_roots=guard, precond=args.precond, graph=args.graph,
                         select_rule=args.select_rule, max_subspace=msub)
        torch.cuda.synchronize()
        times.append(time.time() - ts)
        durations.append(times[-1])
        print(f"[bench] mixed step {s}: {times[-1]:.2f} s", file=sys.stderr, flush=True)
        if not out["converged"]:
            raise SystemExit("mixed-precision Davidson did not converge")
        dmax = max(dmax, float(np.abs(out["evals"] - ref["evals"]).max()))
        last = out
    wall = time.time() - t_all
    if dmax > 1e-8:
        raise SystemExit(f"mixed-precision roots differ from FP64 by {dmax:.3e} Eh (> 1e-8)")
    mixed_s = float(np.mean(times))

    # dominant kernel: the FP32 sigma over the block of the FP32 phase (the
    # roots plus the guard roots); the 16-column FP32 and FP64 launches beside it
    n = H.n
    reps = args.sigma_reps if not direct else 2
    warm = 3 if not direct else 1
    kr = k + guard
    ms32 = time_sigma(H, api, torch, k, api.FP32, reps, mode, warm)
    ms64 = time_sigma(H, api, torch, k, api.FP64, reps, mode, warm)
    msr = time_sigma(H, api, torch, kr, api.FP32, reps, mode, warm) if kr != k else ms32
    if direct:
        nbf = cfg.nbf
        n2 = (nbf * (nbf + 1) // 2) * (nbf * (nbf + 1) // 2 + 1) // 2
        bytes_r = kr * n * (4 + 4) + 16 * cfg.ncfg + 8 * n2
        bytes64 = k * n * (8 + 8) + 16 * cfg.ncfg + 8 * n2
        kernel = (f"k_direct<float,{kr},true,false> (direct FP32 sigma, {kr} vectors: "
                  f"{k} roots + {guard} guard roots)")
        if engine == 2:
            t32 = H.direct_ras_timing(api.FP32, kr, 3)
            t64 = H.direct_ras_timing(api.FP64, k, 3)
    else:
        bytes_r = nnz * (4 + 4) + (n + 1) * 8 + kr * n * (4 + 4)
        bytes64 = nnz * (8 + 4) + (n + 1) * 8 + k * n * (8 + 8)
        kernel = f"k_spmm<float,{kr}> (stored FP32 sigma, {kr} vectors)"
    ach32 = bytes_r / (msr * 1e-3) / 1e9
    ach64 = bytes64 / (ms64 * 1e-3) / 1e9
    traffic = measured_traffic(args.workload, kr)

    hbuild = hbuild_measure(api, torch, not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        if direct:
            cpu = cpu_direct(args, cfg, ints, k, last["stats"]["sigma_calls"])
        else:
            cpu = cpu_stored(args, cfg, ints, k, ref)

    if direct and engine == 2:
        sig_desc = (f"direct sigma, string-driven engine ({ras['ndet']} determinants in {ras['sectors']} spin "
                    f"sectors, {ras['blocks']} MFMA products, {ras['flops_per_column']:.4g} flops per column)")
    elif direct:
        sig_desc = f"direct sigma ({npairs} ordered configuration pairs, {nelems} structural elements per sigma)"
    else:
        sig_desc = f"stored H nnz_lower={nnz}"
    desc = (f"{args.workload}: PSOCI SEF space n={n}, ncfg={cfg.ncfg}, " + sig_desc
            + f", {k} roots, mixed-precision block Davidson, residual tol {args.tol:g}")
    line = {
        "metric": METRIC,
        "value": round(mixed_s, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": warm_done,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "wall_budget_s": args.budget,
        "ms_per_step": round(mixed_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 sigma and f32 basis -> f64 sigma and basis (mixed); f64 subspace matrices and eigensolves",
        "data": ("synthetic (seeded integrals and configuration list, written to a SIFS-style file and a "
                 "configuration file and read back by the library)"),
        "config": {"workload": desc, "roots": k, "n_sef": n, "n_cfg": cfg.ncfg,
                   "sigma": "direct" if direct else "stored"},
        "fp64_seconds": round(fp64_s, 6),
        "speedup_vs_fp64": round(fp64_s / mixed_s, 3),
        "max_root_diff_eh": dmax,
        "root_diffs_eh": [float(x) for x in np.abs(last["evals"] - ref["evals"])],
        "iterations_mixed": last["stats"]["iterations"],
        "guard_roots_fp32_phase": guard,
        "max_subspace": msub if msub else 5 * k,
        "preconditioner": ("diagonal (the reference's)" if args.precond == 0 else
                           "diagonal with Olsen's correction"),
        "select_rule": ("padded-width fill" if args.select_rule == 0 else "the reference's (unconverged roots)"),
        "switch_iter": last["stats"]["switch_iter"],
        "iterations_fp64": ref["stats"]["iterations"],
        "sigma_calls_mixed": last["stats"]["sigma_calls"],
        "host_eig_seconds_mixed": round(last["stats"]["eig_seconds"], 6),
        "graph_launches_mixed": last["stats"]["graph_launches"],
        "sigma_seconds_mixed": round(last["stats"]["sigma_seconds"], 6),
        "setup_seconds": round(setup_s, 4),
        "sigma_fp32_ms": round(ms32, 5),
        "sigma_fp64_ms": round(ms64, 5),
        f"sigma_fp32_ms_k{kr}": round(msr, 5),
        "sigma_fp64_gbs": round(ach64, 1),
        "roofline": {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach32 / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_launch"] if traffic else None, "kernel": kernel,
                     "algorithmic_bytes_per_launch": int(bytes_r), "launch_ms": round(msr, 4)},
        "cpu_baseline": cpu,
        "hbuild": hbuild,
        "wall_seconds_timed": round(wall, 3),
    }
    if traffic:
        line["roofline"]["traffic_source"] = traffic["source"]
        line["roofline"]["traffic_gbs"] = round(traffic["bytes_per_launch"] / (msr * 1e-3) / 1e9, 1)
    if not direct:
        line["stored_shape"] = {"fp32": H.stored_shape(api.FP32), "fp64": H.stored_shape(api.FP64)}
    if direct and engine == 2:
        g32_ms, g32_fl = sum(t32["gemm_ms"]), sum(t32["gemm_flops"])
        g64_ms, g64_fl = sum(t64["gemm_ms"]), sum(t64["gemm_flops"])
        ach = g32_fl / (g32_ms * 1e-3) / 1e12
        ach64f = g64_fl / (g64_ms * 1e-3) / 1e12
        traffic = measured_traffic(args.workload, f"ras{kr}")
        line["roofline"] = {
            "bound": "mfma", "achieved": round(ach, 2), "peak": MFMA_PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
            "frac": round(ach / MFMA_PEAK_TFLOPS["fp32"], 4),
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": (f"k_ras_gemm_lds<float,16,*> (the string-driven sigma's dense products: all "
                       f"{ras['blocks']} launches of one {kr}-column FP32 sigma)"),
            "algorithmic_flops_per_sigma": g32_fl, "gemm_ms_per_sigma": round(g32_ms, 4),
            "fp64": {"achieved": round(ach64f, 2), "peak": MFMA_PEAK_TFLOPS["fp64"],
                     "frac": round(ach64f / MFMA_PEAK_TFLOPS["fp64"], 4), "gemm_ms_per_sigma": round(g64_ms, 4),
                     "algorithmic_flops_per_sigma": g64_fl}}
        if traffic:
            line["roofline"]["traffic_source"] = traffic["source"]
        line["direct"] = {
            "engine": "string-driven (resolutions of the identity over N-1 / N-2 strings, MFMA products)",
            "plan": ras,
            "timing_fp32": {kk: (round(v, 4) if isinstance(v, float) else [round(x, 4) for x in v])
                            for kk, v in t32.items() if kk != "gemm_flops"},
            "timing_fp64": {kk: (round(v, 4) if isinstance(v, float) else [round(x, 4) for x in v])
                            for kk, v in t64.items() if kk != "gemm_flops"},
            "sigma_hbm_gbs_algorithmic": {"fp32": round(ach32, 1), "fp64": round(ach64, 1)},
            "note": ("time per sigma call by launch group: SEF -> determinant scatter, dense products by kind "
                     "(alpha-beta, alpha-alpha, beta-beta), SEF gather; the products run on MFMA, so the "
                     "roofline is the FP32 MFMA peak")}
    elif direct:
        dp = H.direct_det_pairs()
        line["direct"] = {"ordered_configuration_pairs": npairs, "structural_sef_elements_per_sigma": nelems,
                          "connected_determinant_pairs_per_sigma": dp,
                          "schedule": {"fp32": H.direct_schedule(api.FP32), "fp64": H.direct_schedule(api.FP64)},
                          f"det_pairs_per_s_fp32_k{k}": round(dp / (ms32 * 1e-3), 1),
                          f"det_pairs_per_s_fp32_k{kr}": round(dp / (msr * 1e-3), 1),
                          f"det_pairs_per_s_fp64_k{k}": round(dp / (ms64 * 1e-3), 1),
                          "note": ("the work unit is the connected determinant pair (the structural SEF elements "
                                   "are ~16x more, most of them zero and never generated)")}
        line["roofline"]["limiter"] = (
            "not the algorithmic bytes: element generation on the VALU (84 % of the issue rate at 16 FP32 columns) "
            "and x-row gathers that miss L2 (24 FP32 and all FP64 launches; fabric traffic 4.7-5.5 TB/s), "
            "profiles/k_direct_mrci10x24_counters.md")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
