"""The command-line driver psc_run on the GPU: its text outputs (the
reference's file names and columns) against the oracle run with the same
parameters, for the fixed-step cadence and for the reference's adaptive
time-step loop (pspectcosmo.cpp:601-697)."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests._util import parity_setup

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PSC_RUN = os.path.join(ROOT, "pspectcosmo_amd", "psc_run")

COMMON = ["N=16", "kIR=4", "g2=2000", "omegaStar=0.02", "fStar=1", "MPl=1", "initfield=1,0",
          "initderivs=-0.7,0", "seed=1", "precision=fp64", "integrator=VV"]


def run_driver(tmp_path, *args):
    print("psc_run", *args, flush=True)
    r = subprocess.run([PSC_RUN, *COMMON, *args, f"dir={tmp_path}"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def rows(path):
    out = []
    for line in open(path):
        line = line.strip()
        out.append([float(x) for x in line.split()] if line else None)
    return out


def oracle(dt):
    cfg, model, ic = parity_setup(N=16, kIR=4.0, dt=dt, g2=2e3, omega_star=2e-2, f0=(1.0, 0.0),
                                  fd0=(-0.7, 0.0), seed=1)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    return cfg, o


def oracle_means(o, alpha=0):
    f, p = o.fields()
    a = o.scalars()[1]
    aop = a ** (alpha - 3)
    out = []
    for q in range(f.shape[0]):
        fm, pm = f[q].mean(), p[q].mean()
        f2, p2 = (f[q] ** 2).mean(), (p[q] ** 2).mean()
        out.append([fm, aop * pm, f2, aop ** 2 * p2, math.sqrt(max(f2 - fm ** 2, 0)),
                    aop * math.sqrt(max(p2 - pm ** 2, 0))])
    return out


def close(a, b, rtol=2e-6, atol=1e-300):
    np.testing.assert_allclose(np.asarray(a, float), np.asarray(b, float), rtol=rtol, atol=atol)


def test_fixed_step_outputs_match_oracle(tmp_path):
    run_driver(tmp_path, "dt=0.01", "nsteps=10", "freq_steps=5", "infreq_steps=10")
    cfg, o = oracle(0.01)
    energies = rows(tmp_path / "average_energies.txt")
    means = [rows(tmp_path / f"average_scalar_{q}.txt") for q in range(2)]
    scale = rows(tmp_path / "average_scale_factor.txt")
    for r in range(2):
        o.step(5)
        e = o.energies()
        t = o.scalars()[0]
        assert abs(energies[r][0] - t) < 1e-3
        # K0 G0 K1 G1 V0 V1 total (energy_density, output.cpp:57-115)
        close(energies[r][1:], [e[0], e[1], e[2], e[3], e[4], e[5], e[6]])
        for q, m in enumerate(oracle_means(o)):
            close(means[q][r][1:], m, rtol=2e-6, atol=1e-12)
        s = o.scalars()
        close(scale[r][1:3], s[1:3], rtol=0, atol=2e-6)   # fixed 6 decimals
    for q in range(2):
        spec = [x for x in rows(tmp_path / f"spectra_scalar_{q}.txt") if x is not None]
        ref = o.spectra(q)
        assert len(spec) == len(ref)
        close(np.array(spec)[:, 1:4], ref[:, 1:4], rtol=2e-6, atol=1e-300)
        assert np.array_equal(np.array(spec)[:, 4], ref[:, 4])
        ev = rows(tmp_path / f"spectra_scalar_{q}_evolution.txt")
        assert len(ev[0]) == len(ref) + 1 and len(ev[1]) == len(ref) + 1
    times = rows(tmp_path / "average_spectra_times.txt")
    assert [x[0] for x in times] == [0.1]
    assert "Steps=10" in open(tmp_path / "info.txt").read()


def test_adaptive_step_loop_matches_oracle(tmp_path):
    tf, tfreq, tinf = 0.03, 0.01, 0.03
    run_driver(tmp_path, "dt=auto", f"tf={tf}", f"tOutputFreq={tfreq}", f"tOutputInfreq={tinf}")
    cfg, o = oracle(0.01)
    kmax = 4.0 * 16 * math.sqrt(3.0) / 2.0
    t, a = 0.0, 1.0
    next_freq, next_inf = tfreq, tinf
    steps = 0
    expect = []
    while t < tf:
        dt = min(2.0 / (kmax / a) / 10, tfreq)
        if t + dt > next_freq or t + dt > next_inf:
            dt = min(next_freq - t, next_inf - t)
        o.set_dt(dt)
        o.step(1)
        steps += 1
        s = o.scalars()
        t, a = s[0], s[1]
        if abs(t - next_freq) < 1e-10:
            next_freq += tfreq
            expect.append((t, o.energies()[:7]))
        if abs(t - next_inf) < 1e-10:
            next_inf += tinf
    energies = rows(tmp_path / "average_energies.txt")
    assert len(energies) == len(expect) == 3
    for row, (t, e) in zip(energies, expect):
        assert abs(row[0] - t) < 1e-3
        close(row[1:], e)
    assert f"Steps={steps}" in open(tmp_path / "info.txt").read()


def test_checkpoint_and_continue_match_an_uninterrupted_run(tmp_path):
    a_dir, b_dir = tmp_path / "a", tmp_path / "b"
    a_dir.mkdir()
    b_dir.mkdir()
    common = ["dt=0.01", "freq_steps=5", "infreq_steps=5"]
    run_driver(a_dir, *common, "nsteps=10")
    run_driver(b_dir, *common, "nsteps=5", "checkpoint=1")
    assert (b_dir / "grid.img").stat().st_size == 4 + 3 * 8 + 2 * 2 * 16 ** 3 * 8
    r = run_driver(b_dir, *common, "nsteps=10", "checkpoint=1", "continue_run=1")
    assert "Resuming run at t = 0.05" in r.stdout
    ea, eb = rows(a_dir / "average_energies.txt"), rows(b_dir / "average_energies.txt")
    assert len(ea) == len(eb) == 2
    close(eb[1], ea[1], rtol=2e-6)
    sa = [x for x in rows(a_dir / "spectra_scalar_1.txt") if x is not None]
    sb = [x for x in rows(b_dir / "spectra_scalar_1.txt") if x is not None]
    close(sb, sa, rtol=2e-6)


def test_device_initial_conditions_give_the_same_run(tmp_path):
    a_dir, b_dir = tmp_path / "host", tmp_path / "device"
    a_dir.mkdir()
    b_dir.mkdir()
    common = ["dt=0.01", "nsteps=10", "freq_steps=5", "infreq_steps=10"]
    run_driver(a_dir, *common)
    run_driver(b_dir, *common, "ic=device")
    for name in ("average_energies.txt", "average_scalar_0.txt", "average_scalar_1.txt", "spectra_scalar_1.txt"):
        ra = [x for x in rows(a_dir / name) if x is not None]
        rb = [x for x in rows(b_dir / name) if x is not None]
        close(rb, ra, rtol=2e-6, atol=1e-300)


def test_padded_run_matches_oracle(tmp_path):
    run_driver(tmp_path, "dt=0.01", "nsteps=10", "freq_steps=5", "infreq_steps=10", "pad=2")
    cfg, model, ic = parity_setup(N=16, kIR=4.0, dt=0.01, g2=2e3, omega_star=2e-2, f0=(1.0, 0.0),
                                  fd0=(-0.7, 0.0), seed=1, pad=2)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, sc[0], sc[1], 0.0)
    o.load(fk, pk, 0.0, sc[0], sc[1], o.kernel_a())
    energies = rows(tmp_path / "average_energies.txt")
    for r in range(2):
        o.step(5)
        e = o.energies()
        close(energies[r][1:], [e[0], e[1], e[2], e[3], e[4], e[5], e[6]])
    spec = [x for x in rows(tmp_path / "spectra_scalar_1.txt") if x is not None]
    close(np.array(spec)[:, 1:4], o.spectra(1)[:, 1:4], rtol=2e-6, atol=1e-300)
