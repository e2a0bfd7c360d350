"""Throughput of the spectral-viscosity Euler SSP-RK3 step on one MI355X.

Default workload: 3D Taylor-Green, N = 512^3 (M = 768 dealiasing grid), FP64 and
FP32-mixed, C = 0.2, SmoothCutoff1D(eps = 0.05/N, s = 1, k0 = 2 pi N/4).  A "step"
is one SSP-RK3 step (three RHS evaluations) with dt taken on the device.

Prints one JSON line: value = grid points x RK steps per second (Mpts*RK-steps/s)
of the FP64 path, the FP32-mixed figures beside it, the roofline of the dominant
kernel family (algorithmic bytes / hipEvent-timed launch duration vs 8 TB/s),
and the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
AUTOTUNE = True
METRIC = "Mpts·RK-steps/s at 512^3 FP64 vs FP32-mixed; ensemble samples/s; p99 latency"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def tg3d_device(N: int):
    """u = (sin x cos y cos z, -cos x sin y cos z, 0), x = 2 pi i / N, built on the GPU."""
    import torch

    x = 2 * math.pi * torch.arange(N, dtype=torch.float64, device="cuda") / N
    s, c = torch.sin(x), torch.cos(x)
    u = torch.empty((3, N, N, N), dtype=torch.float64, device="cuda")
    u[0] = s[:, None, None] * c[None, :, None] * c[None, None, :]
    u[1] = -c[:, None, None] * s[None, :, None] * c[None, None, :]
    u[2] = 0
    return u


def dsl_device(N: int, seed: int = 1):
    import torch
    from azeban_amd import init as I

    return torch.from_numpy(I.double_shear_layer(N, seed=seed)).to("cuda")


# kernel of each family at the 512^3 FP64 workload, as named in the counter summary
_FAMILY_KERNEL = {"row_fused": "azb::euler_row6_kernel<double>", "kspace_rk": "azb::kspace_kernel<double, 3>"}
COUNTERS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "counters_tg3d_512.json")


def counted_traffic(dim: int, N: int, family: str):
    """HBM bytes per launch of the family's kernel from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes of rocprofv3, scaled by a copy of known size);
    None when no summary for this workload and kernel exists."""
    if (dim, N) != (3, 512) or family not in _FAMILY_KERNEL or not os.path.exists(COUNTERS_FILE):
        return None
    with open(COUNTERS_FILE) as f:
        summ = json.load(f)
    for name, rec in summ["kernels"].items():
        if name.startswith(_FAMILY_KERNEL[family]):
            return round(rec["hbm_bytes"], 0)
    return None


def measured_copy_bandwidth(gib: float = 4.0, reps: int = 10) -> float:
    """Achievable HBM bandwidth: the library's streaming copy (read + write) of `gib` GiB."""
    import ctypes as C

    from azeban_amd._capi import check, load

    out = C.c_double()
    check(load().azb_copy_bandwidth(gib, reps, C.byref(out)))
    return out.value


def run_precision(A, dim, N, u_hat0, precision, steps, warmup):
    import torch

    g = A.Grid(dim, N)
    visc = A.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4)
    eq = A.IncompressibleEuler(g, visc, precision)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(dim), eq), u_hat0)
    tuning = None
    if AUTOTUNE:
        log(f"{precision}: autotune")
        tuning = eq.autotune(sim.u(), reps=3)
    t_far = 1e9
    log(f"{precision}: warmup {warmup} steps")
    if warmup:
        sim.simulate_until(t_far, max_steps=warmup, chunk=warmup)
    torch.cuda.synchronize()
    log(f"{precision}: timing {steps} steps")
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    n = sim.simulate_until(t_far, max_steps=steps, chunk=steps)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    assert n == steps, (n, steps)
    # second pass with per-family hipEvents for the roofline
    log(f"{precision}: profiled pass {steps} steps")
    eq.profile_reset()
    eq.profile_enable(True)
    sim.simulate_until(t_far, max_steps=steps, chunk=steps)
    fam = eq.profile_read()
    eq.profile_enable(False)
    return {"eq": eq, "sim": sim, "wall_s": wall, "gpu_ms": gpu_ms, "families": fam, "tuning": tuning}


def cpu_info() -> dict:
    """The host the CPU baseline runs on: CPU model, usable cores, and whether a system
    FFTW library is present (the reference's CPU FFT)."""
    import subprocess

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    try:
        res = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True, timeout=30)
        fftw = sorted({ln.split()[0] for ln in res.stdout.splitlines() if "fftw3" in ln})
    except (OSError, subprocess.SubprocessError):
        fftw = []
    return {"cpu_model": model, "cores": cores, "fftw3_libraries": fftw or "none (ldconfig -p lists no fftw3)"}


def _oracle_setup(dim: int, Ns: int):
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import azeban_oracle as O
    from azeban_amd import init as I

    if dim == 3:
        u0 = O.initialize(I.taylor_green_3d(Ns))
    elif Ns <= 128:
        u0 = O.initialize(I.taylor_green_2d(Ns))
    else:
        u0 = O.initialize(I.double_shear_layer(Ns))
    return O, u0


def time_oracle(dim: int, Ns: int, workers, threads: int, steps: int = 0, t_end: float = 0.0, runs: int = 3,
                u0=None):
    """Median over `runs` of the oracle's seconds per SSP-RK3 step (either `steps` steps or
    to `t_end`), with FFTs on `workers` threads (None: numpy's single-threaded pocketfft)
    and compute_B / compute_dudt on `threads` threads.  Returns (record, last simulation)."""
    O, u_init = _oracle_setup(dim, Ns)
    u0 = u_init if u0 is None else u0
    per_step, sim = [], None
    for run in range(runs):
        eq = O.IncompressibleEuler(O.Grid(dim, Ns), O.baseline_visc(Ns), workers=workers, threads=threads)
        sim = O.Simulation(eq, u0, 0.2)
        t0 = time.perf_counter()
        if t_end > 0:
            sim.simulate_until(t_end)
            n = len(sim.dts)
        else:
            for _ in range(steps):
                sim.step(1e9)
            n = steps
        el = time.perf_counter() - t0
        per_step.append(el / n)
        log(f"cpu {dim}D N={Ns} workers={workers} threads={threads} run {run}: {n} steps, {el:.2f} s")
    med = sorted(per_step)[len(per_step) // 2]
    return {"s_per_step": med, "Mpts_RK_steps_per_s": Ns**dim / med / 1e6, "steps_per_run": n, "runs": runs,
            "fft_threads": workers or 1, "loop_threads": threads}, sim


def cpu_baseline(dim: int, runs: int = 3):
    """The FP64 oracle on a bounded sample of the workload (3D Taylor-Green 128^3 or the 2D
    double shear layer 1024^2), median of `runs` runs, in three threadings: the
    reference's CPU path (single-threaded FFT, compute_B / compute_dudt on every core, as
    its OpenMP loops), one thread, and every stage on every core.  The headline value is
    the reference threading."""
    info = cpu_info()
    ncores = info["cores"]
    Ns = 128 if dim == 3 else 1024
    steps = 2 if dim == 3 else 5
    ref, sim = time_oracle(dim, Ns, None, ncores, steps=steps, runs=runs)
    one, _ = time_oracle(dim, Ns, None, 1, steps=steps, runs=runs)
    allc, _ = time_oracle(dim, Ns, ncores, ncores, steps=steps, runs=runs)
    name = f"3D Taylor-Green {Ns}^3" if dim == 3 else f"2D double shear layer {Ns}^2"
    O, u0 = _oracle_setup(dim, Ns)
    return {
        "value": ref["Mpts_RK_steps_per_s"],
        "unit": "Mpts*RK-steps/s",
        "cores": ncores,
        "kind": "port",
        "sample": (f"{name}, {steps} SSP-RK3 steps per run, median of {runs} runs, FP64 numpy oracle with "
                   f"the reference's CPU threading: FFT single-threaded (numpy pocketfft, not FFTW), "
                   f"compute_B / compute_dudt on {ncores} threads"),
        "host": info,
        "reference_threading": ref,
        "one_thread": one,
        "all_threads": allc,
        "gpu_fp64_vs_oracle": _gpu_vs_oracle(dim, Ns, u0, sim, steps),
    }


def run_cpu_plan(args):
    """The CPU-baseline plan in full, on this host: 2D Taylor-Green 128^2 to t = 1 and the
    2D double shear layer 1024^2 for 100 steps with the reference's threading (and the
    former on one thread), and one 3D Taylor-Green 512^3 step per run with every stage on
    every core (a single-threaded 768^3 FFT takes minutes per transform).  Median of
    `--cpu-runs` runs each."""
    info = cpu_info()
    nc = info["cores"]
    runs = args.cpu_runs
    out = {"metric": "CPU baseline (FP64 oracle)", "unit": "Mpts*RK-steps/s", "host": info, "cases": {}}
    log("cpu plan: 2D TG 128^2 to t=1")
    out["cases"]["tg2d_128_t1_reference_threading"], _ = time_oracle(2, 128, None, nc, t_end=1.0, runs=runs)
    out["cases"]["tg2d_128_t1_one_thread"], _ = time_oracle(2, 128, None, 1, t_end=1.0, runs=runs)
    log("cpu plan: 2D DSL 1024^2, 100 steps")
    out["cases"]["dsl2d_1024_100_steps_reference_threading"], _ = time_oracle(2, 1024, None, nc, steps=100,
                                                                              runs=runs)
    log("cpu plan: 3D TG 512^3")
    out["cases"]["tg3d_512_all_threads"], _ = time_oracle(3, 512, nc, nc, steps=args.cpu_512_steps, runs=runs)
    out["note"] = ("FFT: numpy/scipy pocketfft, not FFTW (no fftw3 library on the host); 'reference "
                   "threading' = single-threaded FFT with compute_B / compute_dudt on every core, as "
                   "the reference's CPU path; the 512^3 case runs its FFTs on every core too")
    print(json.dumps(out), flush=True)


def _gpu_vs_oracle(dim, Ns, u0, osim, steps):
    """The FP64 GPU path over the same sample (same initial state, same steps), checked
    against the oracle's final state: relative L2 over the spectral coefficients."""
    import numpy as np
    import torch

    import azeban_amd as A

    g = A.Grid(dim, Ns)
    eq = A.IncompressibleEuler(g, A.SmoothCutoff1D(0.05 / Ns, 1.0, 2 * math.pi * Ns / 4))
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(dim), eq), torch.from_numpy(u0).to("cuda"))
    n = sim.simulate_until(1e9, max_steps=steps, chunk=steps)
    torch.cuda.synchronize()
    got = sim.u().cpu().numpy()
    ref = osim.u
    return {"steps": int(n), "rel_l2": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
            "time_gpu": sim.time(), "time_oracle": osim.time}


def run_ensemble(args):
    """Batched 2D ensemble: discontinuous vortex patches, 96 @ 128^2, 96 @ 256^2,
    64 @ 512^2, to t = 1, batched (one graph replay per chunk of steps) vs one
    sample at a time through the single-solve path (timed on a subset per size)."""
    import torch

    from azeban_amd import init as I
    from azeban_amd import initialize
    from azeban_amd.ensemble import Ensemble, Group, solve_one

    counts = {128: 96, 256: 96, 512: 64}
    t_end = args.t_end if args.t_end > 0 else 1.0
    out = {"metric": METRIC, "unit": "samples/s", "higher_is_better": True, "n_gpus": 1,
           "data": "synthetic", "config": {"workload": "batched 2D ensemble, discontinuous vortex patch "
                                           "(radius 0.2, 16-mode 0.01 perturbation), 96@128^2 + 96@256^2 + "
                                           f"64@512^2, t={t_end}, C=0.2", "samples": sum(counts.values())}}
    finals = {}
    for precision in (["fp64", "fp32_mixed"] if not args.no_mixed else ["fp64"]):
        log(f"ensemble {precision}: initial data")
        groups = []
        seed = 0
        for N, S in counts.items():
            us = []
            for _ in range(S):
                us.append(initialize(I.vortex_patch(N, seed=seed)))
                seed += 1
            groups.append(Group(2, N, torch.stack(us).contiguous(), precision))
        init = [g.u.clone() for g in groups]
        ens = Ensemble(groups)
        tuning = None
        if AUTOTUNE:
            log(f"ensemble {precision}: autotune")
            tuning = {str(g.N): rec for g, rec in zip(groups, ens.autotune(reps=2))}
        log(f"ensemble {precision}: batched run")
        res = ens.run(t_end, steps_per_launch=args.ens_chunk)
        torch.cuda.synchronize()
        samples = sum(counts.values())
        rec = {"samples_per_s": round(samples / res["wall_s"], 3), "wall_s": round(res["wall_s"], 3),
               "graph_launches": res["launches"],
               "steps_per_size": {str(g.N): max(st) for g, st in zip(groups, res["steps"])},
               "tuning": tuning}
        if precision == "fp64":
            log("ensemble fp64: one-at-a-time subset")
            import math as _m

            from azeban_amd import Grid, IncompressibleEuler, SmoothCutoff1D

            per = args.ens_subset
            t_single = 0.0
            identical = True
            for gi, g in enumerate(groups):
                # one solver per size, launch shapes tuned for a single sample, reused for
                # every solve (as a caller running samples one after another would)
                eq1 = IncompressibleEuler(Grid(2, g.N), SmoothCutoff1D(0.05 / g.N, 1.0, 2 * _m.pi * g.N / 4))
                if AUTOTUNE:
                    eq1.autotune(init[gi][0], reps=2)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for si in range(per):
                    ref, n = solve_one(2, g.N, init[gi][si], t_end, eq=eq1)
                    torch.cuda.synchronize()
                    identical = identical and bool(torch.equal(ref, g.u[si])) and n == res["steps"][gi][si]
                dt = (time.perf_counter() - t0) / per
                t_single += dt * counts[g.N]
                del eq1
            rec["one_at_a_time_samples_per_s"] = round(samples / t_single, 3)
            rec["one_at_a_time_sample"] = (f"{per} samples per size timed (one solver per size, launch shapes "
                                           "tuned for one sample, reused), scaled by the counts")
            rec["speedup_vs_one_at_a_time"] = round(rec["samples_per_s"] / rec["one_at_a_time_samples_per_s"], 3)
            rec["identical_per_sample_results"] = identical
            out["value"] = rec["samples_per_s"]
            out["dtype"] = "f64"
        finals[precision] = [g.u.clone() for g in groups]
        if precision == "fp32_mixed" and "fp64" in finals:
            # every sample at t_end in both precisions: relative L2 over each size group
            rec["rel_l2_vs_fp64_gpu"] = {
                str(g.N): float(torch.linalg.vector_norm(b - a) / torch.linalg.vector_norm(a))
                for g, a, b in zip(groups, finals["fp64"], finals["fp32_mixed"])}
        out[precision] = rec
        ens.close()
    print(json.dumps(out), flush=True)


def _serve_pair(reqs, args, label):
    """The request set through the continuous-batching server and one at a time; both
    runs see the same arrival times."""
    import torch

    from azeban_amd.serving import Server, serve_one_at_a_time

    log(f"serving {label}: server")
    srv = Server(hbm_budget=int(args.hbm_gib * 2**30), max_active=args.max_active, chunk_steps=args.ens_chunk)
    st = srv.run(reqs)
    torch.cuda.synchronize()
    log(f"serving {label}: one at a time")
    base = serve_one_at_a_time(reqs, chunk=args.ens_chunk)
    got = {r.rid: r for r in base.pop("results")}
    identical = all(res.steps == got[res.rid].steps and bool(torch.equal(res.u_hat, got[res.rid].u_hat))
                    for res in srv.results)

    def rec(x):
        return {"requests_per_s": round(x["throughput_per_s"], 4), "latency_p50_s": round(x["latency_p50_s"], 4),
                "latency_p99_s": round(x["latency_p99_s"], 4), "wall_s": round(x["wall_s"], 3),
                "Mpts_RK_steps_per_s": round(x["grid_point_steps"] / x["wall_s"] / 1e6, 2)}

    server = rec(st)
    server.update({"peak_active": st["peak_active"], "chunk_launches": st["chunk_launches"],
                   "peak_solver_GiB": round(st["peak_solver_bytes"] / 2**30, 2)})
    return {"server": server, "one_at_a_time": rec(base),
            "throughput_vs_one_at_a_time": round(st["throughput_per_s"] / base["throughput_per_s"], 3),
            "identical_results": identical}


def run_serving(args):
    """Poisson arrivals of 2D 128^2-1024^2 and 3D 64^3-256^3 solves of fractional-Brownian
    data (H = 0.75) to t = 0.5 through the continuous-batching server and, on the same
    arrivals, one request at a time: at the configured mean rate of 50/s (far above what
    one GPU serves, so latencies are mostly queueing) and at a rate the server sustains
    (latencies then show service time)."""
    from azeban_amd.serving import poisson_requests

    t_end = args.t_end if args.t_end > 0 else 0.5
    log(f"serving: generating {args.requests} requests at {args.rate}/s")
    reqs = poisson_requests(args.requests, args.rate, seed=1, t_end=t_end)
    hot = _serve_pair(reqs, args, f"{args.rate}/s")
    del reqs
    out = {"metric": METRIC, "unit": "requests/s", "higher_is_better": True, "n_gpus": 1, "dtype": "f64",
           "data": "synthetic", "value": hot["server"]["requests_per_s"],
           "latency_p50_s": hot["server"]["latency_p50_s"], "latency_p99_s": hot["server"]["latency_p99_s"],
           "config": {"workload": f"serving: Poisson arrivals mean {args.rate}/s, 2D 128^2-1024^2 and "
                                  f"3D 64^3-256^3, fBm H=0.75, t={t_end}, FP64", "requests": args.requests,
                      "hbm_budget_GiB": args.hbm_gib, "max_active": args.max_active,
                      "chunk_steps": args.ens_chunk},
           f"rate_{args.rate:g}_per_s": hot}
    if args.sustain_requests > 0:
        log(f"serving: {args.sustain_requests} requests at {args.sustain_rate}/s")
        reqs = poisson_requests(args.sustain_requests, args.sustain_rate, seed=2, t_end=t_end)
        out[f"rate_{args.sustain_rate:g}_per_s"] = _serve_pair(reqs, args, f"{args.sustain_rate}/s")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["tg3d", "dsl2d", "ensemble", "serving", "cpu"], default="tg3d")
    ap.add_argument("--t-end", type=float, default=0.0)
    ap.add_argument("--ens-chunk", type=int, default=16)
    ap.add_argument("--ens-subset", type=int, default=8)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=50.0, help="mean Poisson arrival rate of the serving run")
    ap.add_argument("--sustain-requests", type=int, default=16)
    ap.add_argument("--sustain-rate", type=float, default=0.4,
                    help="a rate the server sustains, for latencies without queueing")
    ap.add_argument("--hbm-gib", type=float, default=200.0)
    ap.add_argument("--max-active", type=int, default=64)
    ap.add_argument("--N", type=int, default=0)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--error-steps", type=int, default=200,
                    help="steps after which the FP32-mixed state is compared with FP64 (both continued)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3, help="runs per CPU-baseline case (median)")
    ap.add_argument("--cpu-512-steps", type=int, default=1, help="3D 512^3 steps per run of the CPU plan")
    ap.add_argument("--calib-copy-gib", type=float, default=0.0,
                    help="run one device-to-device copy of this size first (counter calibration)")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    global AUTOTUNE
    AUTOTUNE = not args.no_autotune

    if args.workload == "cpu":
        return run_cpu_plan(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of this path is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    import torch

    import azeban_amd as A
    from azeban_amd import traffic

    if args.calib_copy_gib > 0:
        n = int(args.calib_copy_gib * 2**30) // 8
        xa = torch.ones(n, dtype=torch.float64, device="cuda")
        ya = torch.empty_like(xa)
        torch.cuda.synchronize()
        ya.copy_(xa)
        torch.cuda.synchronize()
        log(f"calibration copy of {n * 8} bytes done")
        del xa, ya

    if args.workload == "ensemble":
        return run_ensemble(args)
    if args.workload == "serving":
        return run_serving(args)

    dim = 3 if args.workload == "tg3d" else 2
    N = args.N or (512 if dim == 3 else 1024)
    log(f"workload {args.workload} N={N}")
    u_phys = tg3d_device(N) if dim == 3 else dsl_device(N)
    u_hat0 = A.initialize(u_phys)
    del u_phys
    torch.cuda.synchronize()

    res = {"fp64": run_precision(A, dim, N, u_hat0, "fp64", args.steps, args.warmup)}
    if not args.no_mixed:
        res["fp32_mixed"] = run_precision(A, dim, N, u_hat0, "fp32_mixed", args.steps, args.warmup)

    pts = float(N) ** dim

    def summarize(r, precision):
        ms_step = r["gpu_ms"] / args.steps
        fam_bytes = traffic.family_bytes(dim, N, precision)
        kern = {}
        for name, (ms, cnt) in r["families"].items():
            if cnt == 0:
                continue
            avg = ms / cnt
            gbs = fam_bytes[name] / (avg * 1e-3) / 1e9
            kern[name] = {"ms_avg": round(avg, 4), "launches": cnt, "GB_per_launch": round(fam_bytes[name] / 1e9, 4),
                          "GB_s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
        return ms_step, kern

    ms64, kern64 = summarize(res["fp64"], "fp64")
    copy_gbs = measured_copy_bandwidth()
    value = pts * args.steps / (res["fp64"]["gpu_ms"] * 1e-3) / 1e6
    dom = max(kern64.items(), key=lambda kv: kv[1]["ms_avg"] * kv[1]["launches"])
    step_gb = traffic.step_bytes(dim, N, "fp64") / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dom[0],
        "achieved": dom[1]["GB_s"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": dom[1]["frac_hbm"],
        "traffic": counted_traffic(dim, N, dom[0]),
        "measured_copy_GB_s": round(copy_gbs, 1),
        "frac_of_measured_copy": round(dom[1]["GB_s"] / copy_gbs, 4),
        "step_algorithmic_GB": round(step_gb, 3),
        "step_GB_s": round(step_gb / (ms64 * 1e-3), 1),
    }
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpts*RK-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": ("3D Taylor-Green" if dim == 3 else "2D double shear layer")
                   + f" {N}^{dim}, FP64, SSP-RK3 C=0.2, SmoothCutoff eps=0.05/N cutoff N/4",
                   "N": N, "dim": dim, "M": 3 * N // 2},
        "roofline": roofline,
        "kernels_fp64": kern64,
        "tuning_fp64": res["fp64"]["tuning"],
    }
    if "fp32_mixed" in res:
        msm, kernm = summarize(res["fp32_mixed"], "fp32_mixed")
        vm = pts * args.steps / (res["fp32_mixed"]["gpu_ms"] * 1e-3) / 1e6
        sa, sb = res["fp64"]["sim"], res["fp32_mixed"]["sim"]
        done = args.warmup + 2 * args.steps
        if args.error_steps > done:
            log(f"continuing both precisions to {args.error_steps} steps for the error")
            sa.simulate_until(1e9, max_steps=args.error_steps - done, chunk=16)
            sb.simulate_until(1e9, max_steps=args.error_steps - done, chunk=16)
        a, b = sa.u(), sb.u()
        err = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(a))
        out["fp32_mixed"] = {"value": round(vm, 2), "ms_per_step": round(msm, 3),
                             "speedup_vs_fp64": round(vm / value, 3),
                             "rel_l2_vs_fp64_gpu": err,
                             "steps_compared": max(done, args.error_steps),
                             "time_fp64": sa.time(), "time_fp32_mixed": sb.time(),
                             "kernels": kernm, "tuning": res["fp32_mixed"]["tuning"]}
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(dim, args.cpu_runs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
