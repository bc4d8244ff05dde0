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
  --workload adjoint128  gradient of the final <chi^2> w.r.t. g^2, phi0 and the
                         initial chi field at 128^3, FP32-mixed velocity Verlet,
                         500 steps, a checkpoint every 25 steps: adjoint/forward
                         time, extra memory, central differences (eps = 1e-3)
  --workload padded256   the dealiased stepper (V' on 512^3) of the 256^3
                         two-field lattice, Yoshida-4, FP32-mixed; with
                         --n-pad 384 on the 3N/2 padded lattice; with --n 512
                         --n-pad 768 the 512^3 lattice on 768^3

The default run adds, under "workloads", a shorter measurement of each other
configuration (gw512, ensemble128, adjoint128, padded256 with N_pad = 512
and 384, padded512 with N_pad = 768), the attainable
copy bandwidth, the stable variant of the headline lattice (kIR = 1, finite
values), the rate of one-step psc_step calls against one call of many steps,
and the reference's adaptive dt rule on the device against fixed dt.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-workloads]
                       [--workload preheat256|gw512|ensemble128|adjoint128|padded256] [--n-pad 384] [--n 512]
                       [--autotune]
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
from pspectcosmo_amd.lattice import Lattice, generate_ic, generate_gw_ic  # noqa: E402

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


def kernel_roofline(lat, steps, moved=None):
    """Per-kernel durations from HIP events on the library's stream; with a
    dict `moved`, also the algorithmic bytes all kernels move per step."""
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
    if moved is not None:
        moved["bytes_per_step"] = sum(v["bytes"] * v["count"] for v in kinds.values()) / steps
    return dom, achieved, avg_s, d["bytes"], per_kernel


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(fk, pk, sc, model, budget_s=12.0):
    """The oracle's FP64 step with the FFTW3 API (libfftw3, else MKL's FFTW3
    interface; planned per call with FFTW_ESTIMATE as the reference does) on
    every host core the process may use; the oracle's own radix-2 FFT when no
    FFTW3 API loads."""
    from oracle import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    O.set_threads(cores)
    backend = O.set_fft_backend(1, cores)
    if backend is None:
        backend = O.set_fft_backend(0)
    cfg = make_config(N, nflds=2, integrator=PSC_Y4, precision=PSC_FP64, kIR=KIR, dt=DT)
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], 0.0)
    add0 = orc.kernel_a()
    orc.load(fk, pk, 0.0, sc[0], sc[1], add0)
    orc.step(1)   # warm-up step (thread pools, page faults), untimed
    n = 0
    t0 = time.perf_counter()
    while True:
        orc.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 50:
            break
    O.set_fft_backend(0)
    return {"value": N ** 3 * n / el, "unit": "site-updates/s", "cores": cores, "kind": "port",
            "fft_backend": backend, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"{n} Yoshida-4 steps of the same 256^3 two-field lattice after one warm-up step, FP64, "
                      f"the oracle's restatement of the reference step with the FFTW3 API ({backend}) on "
                      f"{cores} threads ({el:.1f} s)"}


def canonical_gw_bytes_per_site_step(s_r, n, substeps=3, nflds=2):
    """Scalar byte model plus 3F(5C + R) + 3F R + 72 C per substep for the GW source,
    its transforms and the u kick / drift."""
    C = 2 * s_r * n * n * (n // 2 + 1)
    R = s_r * n ** 3
    per = nflds * (16 * C + 2 * R) + 3 * nflds * (5 * C + R) + 3 * nflds * R + 72 * C
    return substeps * per / n ** 3


def bench_gw512(args):
    n = 512
    model = make_model(1.0, 1e5)
    ic = make_ic(f0=(1.0, 0.0), fd0=(-0.7, 0.0), seed=1)
    cfg = make_config(n, nflds=2, integrator=PSC_Y4, precision=PSC_FP32, kIR=KIR, dt=0.001,
                      with_gw=1, use_graph=0 if args.no_graph else 1, autotune=1 if args.autotune else 0)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    t0 = time.perf_counter()
    lat.generate_ic_device(0, ic)     # scalar and GW modes generated on the device
    lat.generate_gw_ic_device(0, ic)
    device_sync(lat)
    print(f"[bench] gw512 ICs ready ({(time.perf_counter() - t0) * 1e3:.0f} ms on the device)", file=sys.stderr,
          flush=True)
    if args.autotune:
        print("[bench] autotune", lat.autotune(2), file=sys.stderr, flush=True)
    steps = args.steps
    lat.step(args.warmup)
    device_sync(lat)
    t0 = time.perf_counter()
    lat.step(steps)
    device_sync(lat)
    el = time.perf_counter() - t0
    moved = {}
    _, _, _, _, per_kernel = kernel_roofline(lat, min(steps, 5), moved)
    v = n ** 3 * steps / el
    b = canonical_gw_bytes_per_site_step(4, n)
    out = {"metric": "site-updates/s at 512^3 with GW, FP32-mixed", "value": v, "unit": "site-updates/s",
           "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": el / steps * 1e3,
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "512^3 two-field lattice + 6 GW components, g^2 = 1e5, kIR = 25, "
                                  "Yoshida-4, dt = 0.001", "shape": lat.shape_name()},
           "canonical_bytes_per_site_step": b, "canonical_rate_over_peak": v * b / (HBM_PEAK_GBS * 1e9),
           "moved_bytes_per_step": moved["bytes_per_step"],
           "moved_over_peak": moved["bytes_per_step"] / (el / steps) / (HBM_PEAK_GBS * 1e9),
           "device_bytes": lat.device_bytes(), "kernels": per_kernel}
    lat.close()
    return out


def bench_padded256(args):
    """The dealiased stepper (pad = 2, N_pad = 512, or args.n_pad = 384 for
    3N/2) on the 256^3 two-field lattice, FP32-mixed, Yoshida-4; args.n picks
    another lattice (512 with N_pad = 768)."""
    n = getattr(args, "n", 256)
    n_pad = getattr(args, "n_pad", 0) or 2 * n
    model = make_model(1.0, G2)
    ic = make_ic(f0=(1.0, 0.0), fd0=(-0.7, 0.0), seed=1)
    cfg = make_config(n, nflds=2, integrator=PSC_Y4, precision=PSC_FP32, kIR=KIR, dt=DT, pad=2, n_pad=n_pad,
                      use_graph=0 if args.no_graph else 1)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.generate_ic_device(0, ic)
    device_sync(lat)
    if args.autotune:   # the shapes of the padded passes
        print("[bench] autotune", lat.autotune(2), file=sys.stderr, flush=True)
    steps = args.steps
    lat.step(args.warmup)
    device_sync(lat)
    t0 = time.perf_counter()
    lat.step(steps)
    device_sync(lat)
    el = time.perf_counter() - t0
    moved = {}
    dom, achieved, avg_s, bytes_, per_kernel = kernel_roofline(lat, min(steps, 10), moved)
    v = n ** 3 * steps / el
    # algorithmic bytes per substep in units of X = N^3 x 8 B, with r = N_pad / N
    # (the padded array is r^3 X, its x planes |x| <= N/2 hold r^2 X): drift 3,
    # padding and z inverse 1 + r^2, y inverse, x pass and y forward 2 r^2 each,
    # z forward r + 1, kick 4; 39 X at r = 2, 26.25 X at r = 3/2
    r = n_pad / n
    b = 3 * (9 + 7 * r * r + r) * 8
    out = {"metric": f"site-updates/s at {n}^3, dealiased (N_pad = {n_pad}), FP32-mixed", "value": v,
           "unit": "site-updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": el / steps * 1e3, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{n}^3 two-field lattice, padded V' on {n_pad}^3, g^2 = 2.5e5, kIR = 25, "
                                  "Yoshida-4", "N_pad": n_pad, "shape": lat.shape_name()},
           "roofline": {"kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "avg_us": avg_s * 1e6, "bytes": bytes_},
           "canonical_bytes_per_site_step": b, "canonical_rate_over_peak": v * b / (HBM_PEAK_GBS * 1e9),
           "moved_bytes_per_step": moved["bytes_per_step"],
           "moved_over_peak": moved["bytes_per_step"] / (el / steps) / (HBM_PEAK_GBS * 1e9),
           "device_bytes": lat.device_bytes(), "kernels": per_kernel}
    lat.close()
    return out


def bench_ensemble128(args):
    n = 128
    with open(os.path.join(ROOT, "pspectcosmo_amd", "workloads", "ensemble_128_64.json")) as fh:
        members = json.load(fh)["members"]
    B = len(members)
    cfg = make_config(n, nflds=2, integrator=2, precision=PSC_FP32, kIR=KIR, dt=0.004, batch=B,
                      use_graph=0 if args.no_graph else 1, autotune=1 if args.autotune else 0)
    lat = Lattice(cfg)
    ics = []
    # every member's initial modes generated on the device from its own seed
    # (one untimed call first: the first launch loads the kernels)
    lat.set_model(0, make_model(1.0, members[0]["g2"]))
    lat.generate_ic_device(0, make_ic(f0=(members[0]["phi0"], 0.0), fd0=(-0.7, 0.0), seed=members[0]["seed"]))
    device_sync(lat)
    t0 = time.perf_counter()
    for m in members:
        model = make_model(1.0, m["g2"])
        ic = make_ic(f0=(m["phi0"], 0.0), fd0=(-0.7, 0.0), seed=m["seed"])
        lat.set_model(m["member"], model)
        lat.generate_ic_device(m["member"], ic)
        ics.append((model, ic))
    device_sync(lat)
    ic_device_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    generate_ic(cfg, ics[0][0], ics[0][1])   # the host generator, one member, for comparison
    ic_host_s = time.perf_counter() - t0
    print(f"[bench] ensemble ICs ready: device {ic_device_s * 1e3:.1f} ms for {B} members, host "
          f"{ic_host_s * 1e3:.1f} ms for one", file=sys.stderr, flush=True)
    if args.autotune:
        print("[bench] autotune", lat.autotune(3), file=sys.stderr, flush=True)
    steps = args.steps
    lat.step(args.warmup)
    device_sync(lat)
    t0 = time.perf_counter()
    lat.step(steps)
    device_sync(lat)
    el = time.perf_counter() - t0
    moved = {}
    _, _, _, _, per_kernel = kernel_roofline(lat, min(steps, 20), moved)
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
    for model, ic in ics:
        one.set_model(0, model)
        one.generate_ic_device(0, ic)
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
           "initial_conditions": {"device_ms_all_members": ic_device_s * 1e3,
                                  "host_ms_one_member": ic_host_s * 1e3},
           "canonical_bytes_per_site_step": b, "canonical_rate_over_peak": v * b / (HBM_PEAK_GBS * 1e9),
           "moved_bytes_per_step": moved["bytes_per_step"],
           "moved_over_peak": moved["bytes_per_step"] / (el / steps) / (HBM_PEAK_GBS * 1e9),
           "kernels": per_kernel}
    one.close()
    return out


ADJ_N, ADJ_STEPS, ADJ_EVERY, ADJ_KIR, ADJ_G2, ADJ_DT, ADJ_OMEGA = 128, 500, 25, 2.0, 1e4, 0.004, 1e-4
PHI0, DPHI0 = 1.0, -0.7


_ADJ_IC = {}


def adjoint_case(precision, phi0=PHI0, g2=ADJ_G2, dchi=None, tuned=False, gw=False, graph=True):
    """Context loaded with the adjoint workload's initial state.  The
    fluctuations are generated once (with g^2 = 1e4) and held fixed, as the
    gradient holds them: g2 changes only the coupling, phi0 sets the DC mode
    and da/dt = a H(phi0), dchi (real space) is added to chi."""
    n = ADJ_N
    cfg = make_config(n, nflds=2, integrator=2, precision=precision, kIR=ADJ_KIR, dt=ADJ_DT,
                      omegaStar=ADJ_OMEGA, use_graph=1 if graph else 0, autotune=1 if tuned else 0,
                      with_gw=1 if gw else 0)
    if not _ADJ_IC:
        ic = make_ic(f0=(PHI0, 0.0), fd0=(DPHI0, 0.0), seed=1)
        _ADJ_IC["ic"] = generate_ic(cfg, make_model(1.0, ADJ_G2), ic)
        cgw = make_config(n, nflds=2, integrator=2, precision=precision, kIR=ADJ_KIR, dt=ADJ_DT,
                          omegaStar=ADJ_OMEGA, with_gw=1)
        _ADJ_IC["gw"] = generate_gw_ic(cgw, make_model(1.0, ADJ_G2), ic)
    fk, pk, _ = _ADJ_IC["ic"]
    fk = fk.copy()
    fk[0, 0, 0, 0] = phi0 * n ** 3
    if dchi is not None:
        fk[1] += np.fft.rfftn(dchi)
    H = math.sqrt((DPHI0 ** 2 / 2 + 0.5 * phi0 ** 2) / 3.0)
    lat = Lattice(cfg)
    lat.set_model(0, make_model(1.0, g2))
    lat.load_ic_k(0, fk, pk, 1.0, H)
    if gw:
        lat.load_gw_k(0, *_ADJ_IC["gw"])
    return lat


def gw_fraction_after(lat, steps):
    lat.step(steps)
    return float(lat.gw_spectrum(1)[1][0])


def chi2_after(lat, steps):
    lat.step(steps)
    f, _ = lat.fields()
    return float((f[1] ** 2).mean())


def bench_adjoint128(args):
    steps, every = ADJ_STEPS, ADJ_EVERY
    tuned = not args.no_autotune
    lat = adjoint_case(PSC_FP32, tuned=tuned, graph=not args.no_graph)
    if tuned:
        print("[bench] autotune", lat.autotune(3), file=sys.stderr, flush=True)
    # plain forward of the same run (graph-captured steps)
    lat.step(args.warmup)
    device_sync(lat)
    lat.close()
    lat = adjoint_case(PSC_FP32, tuned=tuned, graph=not args.no_graph)
    shape = lat.shape_name()
    device_sync(lat)
    t0 = time.perf_counter()
    lat.step(steps)
    device_sync(lat)
    fwd = time.perf_counter() - t0
    lat.close()
    lat = adjoint_case(PSC_FP32, tuned=tuned, graph=not args.no_graph)
    lat.adjoint(steps, every, PHI0, DPHI0, want_field=False)   # warm-up: workspace, captured segments
    t0 = time.perf_counter()
    g = lat.adjoint(steps, every, PHI0, DPHI0)
    wall = time.perf_counter() - t0
    dev = lat.device_bytes()
    lat.close()
    print(f"[bench] adjoint fp32: J {g['J']:.6e} dg2 {g['dg2']:.6e} dphi0 {g['dphi0']:.6e} "
          f"fwd {fwd * 1e3:.1f} ms, adjoint fwd {g['forward_ms']:.1f} + bwd {g['backward_ms']:.1f} ms",
          file=sys.stderr, flush=True)
    if args.no_fd and args.no_gw:
        print(json.dumps({"metric": "adjoint/forward time at 128^3, FP32-mixed", "value": g["backward_ms"] / (fwd * 1e3),
                          "unit": "x", "profiling_run": True}))
        return
    if args.no_fd:   # profiling runs: both objectives in FP32, no reference runs
        lgw = adjoint_case(PSC_FP32, tuned=tuned, gw=True, graph=not args.no_graph)
        lgw.adjoint(steps, every, PHI0, DPHI0, want_field=False, objective=1, projection=1)
        ggw = lgw.adjoint(steps, every, PHI0, DPHI0, objective=1, projection=1)
        lgw.close()
        print(json.dumps({"metric": "adjoint/forward time at 128^3, FP32-mixed", "value": g["backward_ms"] / (fwd * 1e3),
                          "unit": "x", "profiling_run": True, "gw_forward_ms": ggw["forward_ms"],
                          "gw_backward_ms": ggw["backward_ms"]}))
        return
    # FP64 GPU gradient of the same run
    l64 = adjoint_case(PSC_FP64)
    g64 = l64.adjoint(steps, every, PHI0, DPHI0)
    l64.close()
    # central differences of the FP32 forward (eps = 1e-3 relative) and of the FP64 forward
    fd = {}
    for name, prec in (("fp32", PSC_FP32), ("fp64", PSC_FP64)):
        e = 1e-3 * ADJ_G2
        Jp = chi2_after(adjoint_case(prec, g2=ADJ_G2 + e), steps)
        Jm = chi2_after(adjoint_case(prec, g2=ADJ_G2 - e), steps)
        dg2 = (Jp - Jm) / (2 * e)
        e = 1e-3 * PHI0
        Jp = chi2_after(adjoint_case(prec, phi0=PHI0 + e), steps)
        Jm = chi2_after(adjoint_case(prec, phi0=PHI0 - e), steps)
        dphi0 = (Jp - Jm) / (2 * e)
        v = np.random.default_rng(7).standard_normal((ADJ_N,) * 3)
        e = 1e-3 * math.sqrt(g["J"])
        Jp = chi2_after(adjoint_case(prec, dchi=e * v), steps)
        Jm = chi2_after(adjoint_case(prec, dchi=-e * v), steps)
        fd[name] = {"dg2": dg2, "dphi0": dphi0, "dchi_dir": (Jp - Jm) / (2 * e)}
    v = np.random.default_rng(7).standard_normal((ADJ_N,) * 3)
    ad_dir = float((g["dchi"] * v).sum())
    ad64_dir = float((g64["dchi"] * v).sum())

    def rel(a, b):
        return abs(a - b) / abs(b)

    # the GW energy fraction objective (TT projection) on the same run with the six u_ij evolved
    lgw = adjoint_case(PSC_FP32, tuned=tuned, gw=True)
    lgw.step(args.warmup)
    device_sync(lgw)
    lgw.close()
    lgw = adjoint_case(PSC_FP32, tuned=tuned, gw=True)
    device_sync(lgw)
    t0 = time.perf_counter()
    lgw.step(steps)
    device_sync(lgw)
    fwd_gw = time.perf_counter() - t0
    lgw.close()
    lgw = adjoint_case(PSC_FP32, tuned=tuned, gw=True)
    lgw.adjoint(steps, every, PHI0, DPHI0, want_field=False, objective=1, projection=1)
    ggw = lgw.adjoint(steps, every, PHI0, DPHI0, objective=1, projection=1)
    lgw.close()
    l64 = adjoint_case(PSC_FP64, gw=True)
    ggw64 = l64.adjoint(steps, every, PHI0, DPHI0, objective=1, projection=1)
    l64.close()
    print(f"[bench] adjoint gw fp32: J {ggw['J']:.6e} fwd {fwd_gw * 1e3:.1f} ms, adjoint fwd "
          f"{ggw['forward_ms']:.1f} + bwd {ggw['backward_ms']:.1f} ms", file=sys.stderr, flush=True)
    fdgw = {}
    for name, prec in (("fp32", PSC_FP32), ("fp64", PSC_FP64)):
        e = 1e-3 * ADJ_G2
        dg2 = (gw_fraction_after(adjoint_case(prec, g2=ADJ_G2 + e, gw=True), steps) -
               gw_fraction_after(adjoint_case(prec, g2=ADJ_G2 - e, gw=True), steps)) / (2 * e)
        e = 1e-3 * PHI0
        dphi0 = (gw_fraction_after(adjoint_case(prec, phi0=PHI0 + e, gw=True), steps) -
                 gw_fraction_after(adjoint_case(prec, phi0=PHI0 - e, gw=True), steps)) / (2 * e)
        v = np.random.default_rng(7).standard_normal((ADJ_N,) * 3)
        e = 1e-3 * math.sqrt(g["J"])
        dd = (gw_fraction_after(adjoint_case(prec, dchi=e * v, gw=True), steps) -
              gw_fraction_after(adjoint_case(prec, dchi=-e * v, gw=True), steps)) / (2 * e)
        fdgw[name] = {"dg2": dg2, "dphi0": dphi0, "dchi_dir": dd}
    v = np.random.default_rng(7).standard_normal((ADJ_N,) * 3)
    gw_dir, gw64_dir = float((ggw["dchi"] * v).sum()), float((ggw64["dchi"] * v).sum())

    ratio = g["backward_ms"] / (fwd * 1e3)
    out = {"metric": "adjoint/forward time at 128^3, 500 VV steps, checkpoint every 25, FP32-mixed",
           "value": ratio, "unit": "x", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": fwd / steps * 1e3, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "128^3 two-field lattice, g^2 = 1e4, kIR = 2, omegaStar = 1e-4, "
                                  "velocity Verlet dt = 0.004, 500 steps, checkpoint every 25; "
                                  "objective <chi^2> at the end", "shape": shape},
           "forward_ms": fwd * 1e3,
           "adjoint_forward_ms": g["forward_ms"], "adjoint_backward_ms": g["backward_ms"],
           "adjoint_total_over_forward": (g["forward_ms"] + g["backward_ms"]) / (fwd * 1e3),
           "adjoint_wall_s": wall,
           "extra_device_bytes": g["extra_bytes"], "context_device_bytes": dev,
           "J": g["J"],
           "grad_fp32": {"dg2": g["dg2"], "dphi0": g["dphi0"], "dchi_dir": ad_dir},
           "grad_fp64": {"dg2": g64["dg2"], "dphi0": g64["dphi0"], "dchi_dir": ad64_dir},
           "fd": fd,
           "rel_fp32_vs_fp64": {"dg2": rel(g["dg2"], g64["dg2"]), "dphi0": rel(g["dphi0"], g64["dphi0"]),
                                "dchi_field": float(np.linalg.norm(g["dchi"] - g64["dchi"]) /
                                                    np.linalg.norm(g64["dchi"]))},
           "rel_fp32_vs_fd32": {k: rel(a, fd["fp32"][k]) for k, a in
                                (("dg2", g["dg2"]), ("dphi0", g["dphi0"]), ("dchi_dir", ad_dir))},
           "rel_fp64_vs_fd64": {k: rel(a, fd["fp64"][k]) for k, a in
                                (("dg2", g64["dg2"]), ("dphi0", g64["dphi0"]), ("dchi_dir", ad64_dir))},
           "gw_fraction": {
               "objective": "rho_GW / rho at the end, TT projection, six u_ij evolved",
               "J": ggw["J"], "forward_ms": fwd_gw * 1e3, "adjoint_forward_ms": ggw["forward_ms"],
               "adjoint_backward_ms": ggw["backward_ms"],
               "backward_over_forward": ggw["backward_ms"] / (fwd_gw * 1e3),
               "total_over_forward": (ggw["forward_ms"] + ggw["backward_ms"]) / (fwd_gw * 1e3),
               "extra_device_bytes": ggw["extra_bytes"],
               "grad_fp32": {"dg2": ggw["dg2"], "dphi0": ggw["dphi0"], "dchi_dir": gw_dir},
               "grad_fp64": {"dg2": ggw64["dg2"], "dphi0": ggw64["dphi0"], "dchi_dir": gw64_dir},
               "fd": fdgw,
               "rel_fp32_vs_fp64": {"dg2": rel(ggw["dg2"], ggw64["dg2"]), "dphi0": rel(ggw["dphi0"], ggw64["dphi0"]),
                                    "dchi_field": float(np.linalg.norm(ggw["dchi"] - ggw64["dchi"]) /
                                                        np.linalg.norm(ggw64["dchi"]))},
               "rel_fp32_vs_fd32": {k: rel(a, fdgw["fp32"][k]) for k, a in
                                    (("dg2", ggw["dg2"]), ("dphi0", ggw["dphi0"]), ("dchi_dir", gw_dir))},
               "rel_fp64_vs_fd64": {k: rel(a, fdgw["fp64"][k]) for k, a in
                                    (("dg2", ggw64["dg2"]), ("dphi0", ggw64["dphi0"]), ("dchi_dir", gw64_dir))}}}
    print(json.dumps(out))


def bench_copy():
    """Attainable HBM copy bandwidth (pspectcosmo_amd/copy_probe: 1 GiB read and
    1 GiB written per launch, 20 launches)."""
    import subprocess
    exe = os.path.join(ROOT, "pspectcosmo_amd", "copy_probe")
    r = subprocess.run([exe, "1024", "20"], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return {"error": r.stderr.strip()[-200:]}
    d = json.loads(r.stdout.strip().splitlines()[-1])
    d["frac_of_peak"] = d["copy_GBps"] / HBM_PEAK_GBS
    return d


STABLE_KIR = 1.0


def stable_lattice(n, integrator, precision, dt, kIR=STABLE_KIR, dt_rule=0, use_graph=1):
    """The headline model on a lattice with kIR = 1, where dt = 0.002 is
    within the stability bound (k_max dt = 0.44 at 256^3, m_chi dt = 1.0), so
    the fields stay finite: ICs generated on the device."""
    cfg = make_config(n, nflds=2, integrator=integrator, precision=precision, kIR=kIR, dt=dt,
                      use_graph=use_graph, dt_rule=dt_rule, omegaStar=1e-4)
    lat = Lattice(cfg)
    lat.set_model(0, make_model(1.0, G2))
    lat.generate_ic_device(0, make_ic(f0=(1.0, 0.0), fd0=(-0.7, 0.0), seed=1))
    return lat


def timed_steps(lat, n, calls=1):
    """Seconds for n steps in `calls` equal calls of psc_step."""
    device_sync(lat)
    t0 = time.perf_counter()
    for _ in range(calls):
        lat.step(n // calls)
    device_sync(lat)
    return time.perf_counter() - t0


def bench_stable(steps=200):
    """The headline configuration with kIR = 1 (finite values) at 256^3,
    Yoshida-4, dt = 0.002, FP32-mixed and FP64."""
    out = {"workload": "256^3 two-field lattice, g^2 = 2.5e5, kIR = 1, Yoshida-4, dt = 0.002 (k_max dt = 0.44), "
                       "omegaStar = 1e-4, device ICs"}
    for name, prec in (("fp32", PSC_FP32), ("fp64", PSC_FP64)):
        lat = stable_lattice(N, PSC_Y4, prec, DT)
        lat.step(5)
        el = timed_steps(lat, steps)
        e = lat.energies(0)
        out[name] = {"value": N ** 3 * steps / el, "ms_per_step": el / steps * 1e3,
                     "finite": bool(np.isfinite(e).all()), "a": float(e[8]),
                     "constraint_violation": float(e[7])}
        lat.close()
    out["fp32_over_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    return out


def bench_single_steps():
    """psc_step(1) called n times against one psc_step(n): the look-ahead
    kick carries across calls, so the one-step calls cost one extra write of
    the synchronised momenta per call (FP32-mixed, stable variant)."""
    out = {}
    for n_lat, n in ((128, 200), (256, 60)):
        for integ, name in ((2, "vv"), (PSC_Y4, "y4")):
            lat = stable_lattice(n_lat, integ, PSC_FP32, DT)
            lat.step(3)
            one = timed_steps(lat, n, 1)
            single = timed_steps(lat, n, n)
            one = min(one, timed_steps(lat, n, 1))
            single = min(single, timed_steps(lat, n, n))
            lat.close()
            out[f"{n_lat}_{name}"] = {"steps": n, "ms_per_step_one_call": one / n * 1e3,
                                      "ms_per_step_single_calls": single / n * 1e3, "ratio": single / one}
    return out


def bench_adaptive(intervals=10, per_interval=20):
    """The reference's dt rule (calculate_dt with the tOutputFreq cap and the
    clamp to output times, pspectcosmo.cpp:213-233, 622-627) evaluated on the
    device inside the captured steps, driven the way psc_run drives it (one
    call per output interval, one read of t and a after it), against fixed dt
    with the same cadence: 256^3, FP32-mixed, Yoshida-4, stable variant."""
    kmax = STABLE_KIR * math.sqrt(3.0) * N / 2.0
    dt0 = 2.0 / kmax / 10.0
    cap = per_interval * dt0 * 1.02   # about per_interval steps per output interval
    res = {}
    for rule in (1, 0):
        lat = stable_lattice(N, PSC_Y4, PSC_FP32, dt0, dt_rule=rule)
        t, a, steps_real, calls = 0.0, 1.0, 0, 0
        n_intervals = intervals + 1   # the first interval is a warm-up
        t_start = None
        s0 = 0
        for k in range(n_intervals):
            if k == 1:
                device_sync(lat)
                t_start = time.perf_counter()
                s0 = lat.step_count(0) if rule else steps_real
            if rule:
                t_out = (k + 1) * cap
                lat.set_output_times(t_out, cap)
                while abs(t - t_out) >= 1e-10:
                    dt = min(2.0 / (kmax / a) / 10.0, cap, t_out - t)
                    lat.step(max(1, int(math.ceil((t_out - t) / dt - 1e-9))))
                    calls += 1
                    sc = lat.scalars(0)
                    t, a = sc[0], sc[1]
            else:
                lat.step(per_interval)
                steps_real += per_interval
                lat.scalars(0)
        device_sync(lat)
        el = time.perf_counter() - t_start
        n_real = (lat.step_count(0) if rule else steps_real) - s0
        res["reference_rule" if rule else "fixed"] = {"value": N ** 3 * n_real / el, "steps": n_real,
                                                      "ms_per_step": el / n_real * 1e3}
        if rule:
            res["reference_rule"]["calls"] = calls
        lat.close()
    res["rule_over_fixed"] = res["reference_rule"]["value"] / res["fixed"]["value"]
    return res


def adjoint_ratios(args):
    """Adjoint cost of the adjoint128 workload (500 VV steps at 128^3,
    FP32-mixed) for both objectives: backward / forward (replays and sweep
    against one forward) and total / forward (with the checkpointed forward),
    with a checkpoint every 25 steps (the configuration) and with one segment
    of all 500 steps (checkpoint_every = nsteps: the forward records every
    step on the way, so nothing is replayed; the same gradient bit for bit,
    for the workspace of a 500-step recording)."""
    steps = ADJ_STEPS
    out = {}
    for every, tag in ((ADJ_EVERY, ""), (steps, "_single_segment")):
        for obj, gw in ((0, False), (1, True)):
            lat = adjoint_case(PSC_FP32, gw=gw)
            lat.step(10)
            device_sync(lat)
            lat.close()
            lat = adjoint_case(PSC_FP32, gw=gw)
            fwd = timed_steps(lat, steps)
            lat.close()
            lat = adjoint_case(PSC_FP32, gw=gw)
            kw = dict(objective=obj, projection=1) if gw else {}
            lat.adjoint(steps, every, PHI0, DPHI0, want_field=False, **kw)
            g = lat.adjoint(steps, every, PHI0, DPHI0, want_field=False, **kw)
            lat.close()
            out[("gw_fraction" if gw else "chi2") + tag] = {
                "checkpoint_every": every, "forward_ms": fwd * 1e3, "adjoint_forward_ms": g["forward_ms"],
                "adjoint_backward_ms": g["backward_ms"],
                "backward_over_forward": g["backward_ms"] / (fwd * 1e3),
                "total_over_forward": (g["forward_ms"] + g["backward_ms"]) / (fwd * 1e3),
                "extra_device_bytes": g["extra_bytes"], "J": g["J"], "dJ_dg2": g["dg2"]}
    return out


def run_workloads(args):
    """Shorter runs of every other configuration, each in its own contexts."""
    w = {}
    quick = argparse.Namespace(**vars(args))
    quick.autotune = False
    for name, fn in (("copy_bandwidth", lambda: bench_copy()),
                     ("stable_variant", lambda: bench_stable()),
                     ("single_step_calls", lambda: bench_single_steps()),
                     ("adaptive_dt", lambda: bench_adaptive()),
                     ("gw512", lambda: bench_gw512(argparse.Namespace(**{**vars(quick), "steps": 20, "warmup": 2}))),
                     # the ensemble's launch shape is autotuned like the headline's (a few ms)
                     ("ensemble128", lambda: bench_ensemble128(argparse.Namespace(
                         **{**vars(quick), "steps": 60, "warmup": 3, "seq_steps": 20,
                            "autotune": not args.no_autotune}))),
                     ("padded256", lambda: bench_padded256(argparse.Namespace(**{**vars(quick), "steps": 10, "warmup": 2}))),
                     ("padded256_n384", lambda: bench_padded256(argparse.Namespace(
                         **{**vars(quick), "steps": 10, "warmup": 2, "n_pad": 384}))),
                     ("padded512_n768", lambda: bench_padded256(argparse.Namespace(
                         **{**vars(quick), "steps": 4, "warmup": 1, "n": 512, "n_pad": 768}))),
                     ("adjoint128", lambda: adjoint_ratios(args))):
        t0 = time.perf_counter()
        print(f"[bench] workload {name} ...", file=sys.stderr, flush=True)
        try:
            w[name] = fn()
        except Exception as e:   # a failed sub-measurement is reported, not fatal to the headline
            w[name] = {"error": f"{type(e).__name__}: {e}"}
        if isinstance(w[name], dict):
            w[name]["bench_seconds"] = time.perf_counter() - t0
    return w


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-graph", action="store_true", help="launch kernels one by one")
    ap.add_argument("--fp32-only", action="store_true", help="profiling runs: FP32 path only")
    ap.add_argument("--workload", default="preheat256",
                    choices=["preheat256", "gw512", "ensemble128", "adjoint128", "padded256"])
    ap.add_argument("--autotune", action="store_true", help="time the launch shapes first")
    ap.add_argument("--n-pad", dest="n_pad", type=int, default=0, help="padded256: N_pad (512 or 384)")
    ap.add_argument("--n", type=int, default=256, help="padded256: the lattice (256, or 512 with --n-pad 768)")
    ap.add_argument("--no-autotune", action="store_true", help="headline: keep the default shapes")
    ap.add_argument("--seq-steps", type=int, default=150, help="ensemble: steps per member run alone")
    ap.add_argument("--no-fd", action="store_true", help="adjoint: skip the finite-difference and FP64 runs")
    ap.add_argument("--no-gw", action="store_true", help="adjoint profiling runs: <chi^2> objective only")
    ap.add_argument("--no-workloads", action="store_true", help="headline only (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the lattice on a single GPU; --gpus "
              f"{args.gpus} is not supported (multi-GPU sharding is a design note only).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "gw512":
        return print(json.dumps(bench_gw512(args)))
    if args.workload == "padded256":
        return print(json.dumps(bench_padded256(args)))
    if args.workload == "adjoint128":
        return bench_adjoint128(args)
    if args.workload == "ensemble128":
        return print(json.dumps(bench_ensemble128(args)))

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
    moved32, moved64 = {}, {}
    dom, achieved, avg_s, bpl, per_kernel = kernel_roofline(lat32, min(args.steps, 50), moved32)
    lat32.close()
    s64 = args.fp64_steps or args.steps
    lat64, el64 = run_gpu(PSC_FP64, fk, pk, sc, model, s64, args.warmup, graph, tune)
    shape64 = lat64.shape_name()
    v64 = N ** 3 * s64 / el64
    _, _, _, _, per_kernel64 = kernel_roofline(lat64, min(s64, 20), moved64)
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
        # whole-step rates: the algorithmic bytes the kernels move per step
        # over the step time (an HBM fraction, below 1), and the canonical
        # 16C + 2R count per field over the step time (a rate in canonical
        # bytes: the pipeline moves fewer bytes than it counts, and the
        # memory-side cache serves part of them, so it can exceed 1)
        "step_rates": {
            "moved_bytes_per_step": {"fp32": moved32["bytes_per_step"], "fp64": moved64["bytes_per_step"]},
            "moved_over_peak": {
                "fp32": moved32["bytes_per_step"] / (el32 / args.steps) / (HBM_PEAK_GBS * 1e9),
                "fp64": moved64["bytes_per_step"] / (el64 / s64) / (HBM_PEAK_GBS * 1e9)},
            "canonical_bytes_per_site_step": {"fp32": canonical_bytes_per_site_step(4),
                                              "fp64": canonical_bytes_per_site_step(8)},
            "canonical_rate_over_peak": {
                "fp32": v32 * canonical_bytes_per_site_step(4) / (HBM_PEAK_GBS * 1e9),
                "fp64": v64 * canonical_bytes_per_site_step(8) / (HBM_PEAK_GBS * 1e9)},
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bpl, "avg_launch_us": avg_s * 1e6},
        "kernels_fp32": per_kernel,
        "kernels_fp64": per_kernel64,
    }
    if not args.no_workloads and not args.fp32_only:
        out["workloads"] = run_workloads(args)
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(fk, pk, sc, model)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
