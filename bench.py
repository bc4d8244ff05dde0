"""Throughput of the MI355X lattice time step (one GPU).

Workload: the two-field preheating lattice V = m^2 phi^2/2 + g^2 phi^2 chi^2/2
at 256^3 (m = 1, g^2 = 2.5e5, phi0 = 1, dphi0 = -0.7, kIR = 25, seeded
Gaussian-random-field fluctuations), Yoshida-4 with dt = 0.002.  The headline
value is the FP32-mixed path (FP32 fields/FFTs, FP64 scale factor and
reductions); the FP64 path is measured on the same GPU in the same run.
Inputs are resident in HBM before the timed region.  As specified this
configuration exceeds the leapfrog stability bound at the top lattice momentum
(k_max dt = 11.1), so the fields overflow within a few steps; the bytes moved
and the arithmetic per step do not depend on the values.

Other workloads (own JSON line each):
  --workload gw512       512^3 lattice with the 6 u_ij / pi_u_ij, g^2 = 1e5,
                         Yoshida-4, dt = 0.001, FP32-mixed
  --workload ensemble128 64 independent 128^3 members in one context (seeds,
                         g^2 and phi0 from pspectcosmo_amd/workloads/
                         ensemble_128_64.json), velocity Verlet, dt = 0.004,
                         FP32-mixed, against the same members run one by one

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--workload preheat256|gw512|ensemble128] [--autotune]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pspectcosmo_amd.abi import make_config, make_model, make_ic, PSC_Y4, PSC_FP32, PSC_FP64  # noqa: E402
from pspectcosmo_amd.lattice import Lattice, generate_ic  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "site-updates/s at 256³, FP32-mixed vs FP64 on same GPU; % HBM roofline"
N, KIR, DT, G2 = 256, 25.0, 0.002, 2.5e5


def canonical_bytes_per_site_step(s_r, substeps=3, nflds=2):
    """B_step / N^3 with B_step = s F (16 C + 2 R), C = 2 s_r N^2 (N/2+1), R = s_r N^3."""
    C = 2 * s_r * N * N * (N // 2 + 1)
    R = s_r * N ** 3
    return substeps * nflds * (16 * C + 2 * R) / N ** 3


def device_sync(lat):
    lat.sync()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def run_gpu(precision, fk, pk, sc, model, steps, warmup, graph=True, autotune=False):
    cfg = make_config(N, nflds=2, integrator=PSC_Y4, precision=precision, kIR=KIR, dt=DT,
                      use_graph=1 if graph else 0)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    if autotune:
        rep = lat.autotune(3)
        print(f"[bench] autotune {rep} -> {lat.shape_name()}", file=sys.stderr, flush=True)
    if warmup > 0:
        lat.step(warmup)
    device_sync(lat)
    t0 = time.perf_counter()
    lat.step(steps)
    device_sync(lat)
    t1 = time.perf_counter()
    return lat, t1 - t0


def kernel_roofline(lat, steps):
    """Per-kernel durations from HIP events on the library's stream."""
    lat.reset_kernel_times()
    lat.set_kernel_timing(True)
    lat.step(steps)
    kt = lat.kernel_times()
    lat.set_kernel_timing(False)
    kinds = {k: v for k, v in kt.items() if v["count"] > 0}
    dom = max(kinds, key=lambda k: kinds[k]["ms"])
    d = kinds[dom]
    avg_s = d["ms"] / d["count"] / 1e3
    achieved = d["bytes"] / avg_s / 1e9
    per_kernel = {k: {"avg_us": v["ms"] / v["count"] * 1e3, "count": v["count"],
                      "GBps": v["bytes"] / (v["ms"] / v["count"] / 1e3) / 1e9 if v["ms"] > 0 else 0}
                  for k, v in kinds.items()}
    return dom, achieved, avg_s, d["bytes"], per_kernel


def cpu_baseline(fk, pk, sc, model, budget_s=12.0):
    from oracle import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    O.set_threads(cores)
    cfg = make_config(N, nflds=2, integrator=PSC_Y4, precision=PSC_FP64, kIR=KIR, dt=DT)
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], 0.0)
    add0 = orc.kernel_a()
    orc.load(fk, pk, 0.0, sc[0], sc[1], add0)
    n = 0
    t0 = time.perf_counter()
    while True:
        orc.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 50:
            break
    return {"value": N ** 3 * n / el, "unit": "site-updates/s", "cores": cores, "kind": "port",
            "sample": f"{n} Yoshida-4 steps of the same 256^3 two-field lattice, FP64, oracle "
                      f"restatement with its own radix-2 FFT and OpenMP ({el:.1f} s)"}


def canonical_gw_bytes_per_site_step(s_r, n, substeps=3, nflds=2):
    """Scalar byte model plus 3F(5C + R) + 3F R + 72 C per substep for the GW source,
    its transforms and the u kick / drift."""
    C = 2 * s_r * n * n * (n // 2 + 1)
    R = s_r * n ** 3
    per = nflds * (16 * C + 2 * R) + 3 * nflds * (5 * C + R) + 3 * nflds * R + 72 * C
    return substeps * per / n ** 3


def bench_gw512(args):
    from pspectcosmo_amd.lattice import generate_gw_ic
    n = 512
    model = make_model(1.0, 1e5)
    ic = make_ic(f0=(1.0, 0.0), fd0=(-0.7, 0.0), seed=1)
    cfg = make_config(n, nflds=2, integrator=PSC_Y4, precision=PSC_FP32, kIR=KIR, dt=0.001,
                      with_gw=1, use_graph=0 if args.no_graph else 1, autotune=1 if args.autotune else 0)
    fk, pk, sc = generate_ic(cfg, model, ic)
    u, pu = generate_gw_ic(cfg, model, ic)
    print("[bench] gw512 ICs ready", file=sys.stderr, flush=True)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    lat.load_gw_k(0, u, pu)
    del fk, pk, u, pu
    if args.autotune:
        print("[bench] autotune", lat.autotune(2), file=sys.stderr, flush=True)
    steps = args.steps
    lat.step(args.warmup)
    device_sync(lat)
    t0 = time.perf_counter()
    lat.step(steps)
    device_sync(lat)
    el = time.perf_counter() - t0
    _, _, _, _, per_kernel = kernel_roofline(lat, min(steps, 5))
    v = n ** 3 * steps / el
    b = canonical_gw_bytes_per_site_step(4, n)
    out = {"metric": "site-updates/s at 512^3 with GW, FP32-mixed", "value": v, "unit": "site-updates/s",
           "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": el / steps * 1e3,
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "512^3 two-field lattice + 6 GW components, g^2 = 1e5, kIR = 25, "
                                  "Yoshida-4, dt = 0.001", "shape": lat.shape_name()},
           "canonical_bytes_per_site_step": b, "step_hbm_frac_canonical": v * b / (HBM_PEAK_GBS * 1e9),
           "device_bytes": lat.device_bytes(), "kernels": per_kernel}
    print(json.dumps(out))


def bench_ensemble128(args):
    n = 128
    with open(os.path.join(ROOT, "pspectcosmo_amd", "workloads", "ensemble_128_64.json")) as fh:
        members = json.load(fh)["members"]
    B = len(members)
    cfg = make_config(n, nflds=2, integrator=2, precision=PSC_FP32, kIR=KIR, dt=0.004, batch=B,
                      use_graph=0 if args.no_graph else 1, autotune=1 if args.autotune else 0)
    lat = Lattice(cfg)
    ics = []
    for m in members:
        model = make_model(1.0, m["g2"])
        ic = make_ic(f0=(m["phi0"], 0.0), fd0=(-0.7, 0.0), seed=m["seed"])
        fk, pk, sc = generate_ic(cfg, model, ic)
        lat.set_model(m["member"], model)
        lat.load_ic_k(m["member"], fk, pk, sc[0], sc[1])
        ics.append((model, fk, pk, sc))
    print("[bench] ensemble ICs ready", file=sys.stderr, flush=True)
    if args.autotune:
        print("[bench] autotune", lat.autotune(3), file=sys.stderr, flush=True)
    steps = args.steps
    lat.step(args.warmup)
    device_sync(lat)
    t0 = time.perf_counter()
    lat.step(steps)
    device_sync(lat)
    el = time.perf_counter() - t0
    _, _, _, _, per_kernel = kernel_roofline(lat, min(steps, 20))
    shape = lat.shape_name()
    lat.close()
    v = B * n ** 3 * steps / el
    # the same members one at a time, each in its own single-member context
    seq_steps = args.seq_steps
    cfg1 = make_config(n, nflds=2, integrator=2, precision=PSC_FP32, kIR=KIR, dt=0.004, batch=1,
                       use_graph=0 if args.no_graph else 1)
    one = Lattice(cfg1)
    if args.autotune:
        one.set_shape(shape)
    seq = 0.0
    for model, fk, pk, sc in ics:
        one.set_model(0, model)
        one.load_ic_k(0, fk, pk, sc[0], sc[1])
        one.step(2)
        device_sync(one)
        t0 = time.perf_counter()
        one.step(seq_steps)
        device_sync(one)
        seq += time.perf_counter() - t0
    vseq = B * n ** 3 * seq_steps / seq
    C = 2 * 4 * n * n * (n // 2 + 1)
    R = 4 * n ** 3
    b = 2 * (16 * C + 2 * R) / n ** 3
    out = {"metric": "site-updates/s, 64 x 128^3 batched ensemble, FP32-mixed", "value": v,
           "unit": "site-updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": el / steps * 1e3, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic", "config": {"workload": "64 members of 128^3, seeds 1..64, g^2 log-uniform "
                                                       "[1e4, 1e6], phi0 uniform [0.8, 1.2], VV, dt = 0.004",
                                           "shape": shape},
           "sequential": {"value": vseq, "steps_per_member": seq_steps},
           "batched_over_sequential": v / vseq,
           "canonical_bytes_per_site_step": b, "step_hbm_frac_canonical": v * b / (HBM_PEAK_GBS * 1e9),
           "kernels": per_kernel}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-graph", action="store_true", help="launch kernels one by one")
    ap.add_argument("--fp32-only", action="store_true", help="profiling runs: FP32 path only")
    ap.add_argument("--workload", default="preheat256", choices=["preheat256", "gw512", "ensemble128"])
    ap.add_argument("--autotune", action="store_true", help="time the launch shapes first")
    ap.add_argument("--no-autotune", action="store_true", help="headline: keep the default shapes")
    ap.add_argument("--seq-steps", type=int, default=150, help="ensemble: steps per member run alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the lattice on a single GPU; --gpus "
              f"{args.gpus} is not supported (multi-GPU sharding is a design note only).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "gw512":
        return bench_gw512(args)
    if args.workload == "ensemble128":
        return bench_ensemble128(args)

    model = make_model(1.0, G2)
    ic = make_ic(f0=(1.0, 0.0), fd0=(-0.7, 0.0), seed=1)
    cfg0 = make_config(N, nflds=2, integrator=PSC_Y4, precision=PSC_FP64, kIR=KIR, dt=DT)
    fk, pk, sc = generate_ic(cfg0, model, ic)
    print(f"[bench] ICs ready (256^3, seed 1)", file=sys.stderr, flush=True)

    graph = not args.no_graph
    tune = not args.no_autotune and not args.fp32_only
    lat32, el32 = run_gpu(PSC_FP32, fk, pk, sc, model, args.steps, args.warmup, graph, tune)
    shape32 = lat32.shape_name()
    v32 = N ** 3 * args.steps / el32
    print(f"[bench] fp32: {el32:.3f} s for {args.steps} steps", file=sys.stderr, flush=True)
    if args.fp32_only:
        lat32.close()
        print(json.dumps({"metric": METRIC, "value": v32, "unit": "site-updates/s",
                          "steps": args.steps, "profiling_run": True}))
        return
    dom, achieved, avg_s, bpl, per_kernel = kernel_roofline(lat32, min(args.steps, 50))
    lat32.close()
    s64 = args.fp64_steps or args.steps
    lat64, el64 = run_gpu(PSC_FP64, fk, pk, sc, model, s64, args.warmup, graph, tune)
    shape64 = lat64.shape_name()
    v64 = N ** 3 * s64 / el64
    _, _, _, _, per_kernel64 = kernel_roofline(lat64, min(s64, 20))
    lat64.close()
    print(f"[bench] fp64: {el64:.3f} s for {s64} steps", file=sys.stderr, flush=True)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_256_fp32.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            t = json.load(fh)
        traffic = t.get(dom)

    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "256^3 two-field preheating lattice (V = m^2 phi^2/2 + g^2 phi^2 chi^2/2, "
                        "g^2 = 2.5e5, kIR = 25), Yoshida-4, dt = 0.002, seeded GRF ICs",
            "N": N, "integrator": "yoshida4", "precision": "fp32-mixed",
            "stable_as_specified": False, "shape_fp32": shape32, "shape_fp64": shape64,
        },
        "fp64": {"value": v64, "ms_per_step": el64 / s64 * 1e3, "steps": s64},
        "fp32_over_fp64": v32 / v64,
        "step_hbm_frac_canonical": {
            "fp32": v32 * canonical_bytes_per_site_step(4) / (HBM_PEAK_GBS * 1e9),
            "fp64": v64 * canonical_bytes_per_site_step(8) / (HBM_PEAK_GBS * 1e9),
            "bytes_per_site_step": {"fp32": canonical_bytes_per_site_step(4),
                                    "fp64": canonical_bytes_per_site_step(8)},
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bpl, "avg_launch_us": avg_s * 1e6},
        "kernels_fp32": per_kernel,
        "kernels_fp64": per_kernel64,
    }
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(fk, pk, sc, model)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
