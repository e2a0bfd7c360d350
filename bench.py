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


def cpu_baseline(dim: int, budget_s: float):
    """The FP64 numpy oracle (pocketfft, one thread) on a bounded sample."""
    import numpy as np

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import azeban_oracle as O
    from azeban_amd import init as I

    Ns = 128 if dim == 3 else 1024
    u0 = O.initialize(I.taylor_green_3d(Ns) if dim == 3 else I.double_shear_layer(Ns))
    eq = O.IncompressibleEuler(O.Grid(dim, Ns), O.baseline_visc(Ns))
    sim = O.Simulation(eq, u0, 0.2)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s or steps == 0:
        sim.step(1e9)
        steps += 1
    el = time.perf_counter() - t0
    name = f"3D Taylor-Green {Ns}^3" if dim == 3 else f"2D double shear layer {Ns}^2"
    return {
        "value": Ns**dim * steps / el / 1e6,
        "unit": "Mpts*RK-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{name}, {steps} SSP-RK3 steps in {el:.1f} s, FP64 numpy oracle "
                  f"(pocketfft, not FFTW), 1 thread",
        "gpu_fp64_vs_oracle": _gpu_vs_oracle(dim, Ns, u0, sim, steps),
    }


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


def run_serving(args):
    """Poisson arrivals (mean rate 50/s) of 2D 128^2-1024^2 and 3D 64^3-256^3 solves of
    fractional-Brownian data (H = 0.75) to t = 0.5, continuous batching by HBM footprint."""
    import torch

    from azeban_amd.serving import Server, poisson_requests

    t_end = args.t_end if args.t_end > 0 else 0.5
    log(f"serving: generating {args.requests} requests")
    reqs = poisson_requests(args.requests, 50.0, seed=1, t_end=t_end)
    srv = Server(hbm_budget=int(args.hbm_gib * 2**30), max_active=args.max_active, chunk_steps=args.ens_chunk)
    log("serving: running")
    st = srv.run(reqs)
    torch.cuda.synchronize()
    out = {"metric": METRIC, "unit": "requests/s", "higher_is_better": True, "n_gpus": 1, "dtype": "f64",
           "data": "synthetic", "value": round(st["throughput_per_s"], 3),
           "latency_p50_s": round(st["latency_p50_s"], 4), "latency_p99_s": round(st["latency_p99_s"], 4),
           "wall_s": round(st["wall_s"], 3), "peak_active": st["peak_active"],
           "Mpts_RK_steps_per_s": round(st["grid_point_steps"] / st["wall_s"] / 1e6, 2),
           "config": {"workload": "serving: Poisson arrivals mean 50/s, 2D 128^2-1024^2 and 3D 64^3-256^3, "
                                  f"fBm H=0.75, t={t_end}, FP64", "requests": args.requests,
                      "hbm_budget_GiB": args.hbm_gib, "max_active": args.max_active}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["tg3d", "dsl2d", "ensemble", "serving"], default="tg3d")
    ap.add_argument("--t-end", type=float, default=0.0)
    ap.add_argument("--ens-chunk", type=int, default=16)
    ap.add_argument("--ens-subset", type=int, default=8)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--hbm-gib", type=float, default=200.0)
    ap.add_argument("--max-active", type=int, default=32)
    ap.add_argument("--N", type=int, default=0)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--error-steps", type=int, default=200,
                    help="steps after which the FP32-mixed state is compared with FP64 (both continued)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calib-copy-gib", type=float, default=0.0,
                    help="run one device-to-device copy of this size first (counter calibration)")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    global AUTOTUNE
    AUTOTUNE = not args.no_autotune

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
        out["cpu_baseline"] = cpu_baseline(dim, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
