"""GPU checks at the production lattice sizes and for every launch shape.

N = 128 and 256 are compared with the FP64 oracle directly; at N = 512 the
oracle would need tens of GB of host memory, so the 512-point transforms are
pinned by the exact discrete solution of the free field and by time reversal
of the nonlinear two-field step."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import (PSC_VV2, PSC_Y4, PSC_FP64, PSC_FP32, PSC_EXPANSION_STATIC,
                                 make_config)
from pspectcosmo_amd.lattice import Lattice, generate_ic
from tests._util import parity_setup, rel_err

pytestmark = pytest.mark.gpu


def _gpu(cfg, model, fk, pk, sc):
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    return lat


@pytest.mark.parametrize("N,steps,integrator", [(128, 10, PSC_VV2), (256, 2, PSC_Y4)])
def test_fp64_parity_large(N, steps, integrator):
    cfg, model, ic = parity_setup(N=N, integrator=integrator)
    fk, pk, sc = generate_ic(cfg, model, ic)
    lat = _gpu(cfg, model, fk, pk, sc)
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], 0.0)
    add0 = orc.kernel_a()
    orc.load(fk, pk, 0.0, sc[0], sc[1], add0)
    assert abs(lat.scalars(0)[3] - add0) <= 1e-12 * abs(add0)
    lat.step(steps)
    orc.step(steps)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    e = (rel_err(gk, ok), rel_err(gp, op))
    print(f"N={N}: field {e[0]:.3e} momentum {e[1]:.3e}")
    assert max(e) <= 1e-10
    np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10)


def _free_field_exact(N, precision, steps, tol):
    kIR, m2, dt = 0.5, 1.0, 0.005  # k_max dt = 1.1 at N = 512
    cfg, model, ic = parity_setup(N=N, nflds=1, kIR=kIR, dt=dt, g2=0.0, m2=m2,
                                  precision=precision, expansion=PSC_EXPANSION_STATIC,
                                  omega_star=1e-3, fd0=(0.2, 0.0))
    fk, pk, sc = generate_ic(cfg, model, ic)
    lat = _gpu(cfg, model, fk, pk, (1.0, 0.0))
    lat.step(steps)
    gk, gp = lat.get_k(0)
    i = np.arange(N)
    n = np.where(i <= N // 2, i, i - N).astype(np.float64)
    nz = np.arange(N // 2 + 1, dtype=np.float64)
    w2 = kIR ** 2 * (n[:, None, None] ** 2 + n[None, :, None] ** 2 + nz[None, None, :] ** 2) + m2
    A = 1 - w2 * dt * dt / 2
    C = -w2 * dt * (1 - w2 * dt * dt / 4)
    f, p = fk[0], pk[0]
    for _ in range(steps):
        f, p = A * f + dt * p, C * f + A * p
    ef, ep = rel_err(gk[0], f), rel_err(gp[0], p)
    print(f"N={N} prec={precision} free field: {ef:.3e} {ep:.3e}")
    assert ef < tol and ep < tol


def test_free_field_exact_512_fp64():
    _free_field_exact(512, PSC_FP64, 2, 1e-12)


def test_free_field_exact_512_fp32():
    _free_field_exact(512, PSC_FP32, 2, 2e-6)


def test_time_reversal_512_fp64():
    """Forward then backward steps of the nonlinear expanding two-field
    lattice return to the initial state."""
    N = 512
    cfg, model, ic = parity_setup(N=N, integrator=PSC_VV2)
    fk, pk, sc = generate_ic(cfg, model, ic)
    fwd = _gpu(cfg, model, fk, pk, sc)
    fwd.step(2)
    k1, p1 = fwd.get_k(0)
    s1 = fwd.scalars(0)
    fwd.close()
    cfg_b, _, _ = parity_setup(N=N, integrator=PSC_VV2, dt=-cfg.dt)
    back = Lattice(cfg_b)
    back.set_model(0, model)
    back.load_ic_k(0, k1, p1, s1[1], s1[2])
    back.set_scalars(0, s1[0], s1[1], s1[2], s1[3])
    back.step(2)
    k0, p0 = back.get_k(0)
    s0 = back.scalars(0)
    e = (rel_err(k0, fk), rel_err(p0, pk))
    print("reversal", e, s0)
    assert max(e) < 1e-11
    assert abs(s0[1] - sc[0]) < 1e-12 and abs(s0[2] - sc[1]) < 1e-9 * abs(sc[1])


@pytest.mark.parametrize("precision,N", [(PSC_FP32, 128), (PSC_FP64, 128), (PSC_FP32, 256),
                                         (PSC_FP64, 256), (PSC_FP32, 512), (PSC_FP64, 512)])
def test_every_shape_matches_default(precision, N):
    cfg, model, ic = parity_setup(N=N, integrator=PSC_Y4, precision=precision)
    fk, pk, sc = generate_ic(cfg, model, ic)
    lat = _gpu(cfg, model, fk, pk, sc)
    shapes = lat.shapes()
    assert lat.shape_name() == shapes[0]
    lat.step(2)
    ref = lat.get_k(0)
    lat.close()
    tol = 1e-13 if precision == PSC_FP64 else 2e-6
    for name in shapes[1:]:
        other = Lattice(cfg)
        other.set_shape(name)
        other.set_model(0, model)
        other.load_ic_k(0, fk, pk, sc[0], sc[1])
        other.step(2)
        k = other.get_k(0)
        e = max(rel_err(k[0], ref[0]), rel_err(k[1], ref[1]))
        print(f"{name}: {e:.3e}")
        assert e <= tol, name
        other.close()


def test_autotune_restores_state_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("PSC_TUNE_CACHE", str(tmp_path / "tune.txt"))
    cfg, model, ic = parity_setup(N=128, integrator=PSC_Y4, precision=PSC_FP32, use_graph=1)
    fk, pk, sc = generate_ic(cfg, model, ic)
    lat = _gpu(cfg, model, fk, pk, sc)
    before = lat.get_k(0), lat.scalars(0)
    rep = lat.autotune(2)
    print("autotune", rep, "->", lat.shape_name())
    assert set(rep) == set(lat.shapes())
    after = lat.get_k(0), lat.scalars(0)
    np.testing.assert_array_equal(before[0][0], after[0][0])
    np.testing.assert_array_equal(before[0][1], after[0][1])
    np.testing.assert_array_equal(before[1], after[1])
    text = (tmp_path / "tune.txt").read_text()
    assert lat.shape_name() in text
    cfg2, _, _ = parity_setup(N=128, integrator=PSC_Y4, precision=PSC_FP32, use_graph=1)
    cfg2.autotune = 1
    again = Lattice(cfg2)
    assert again.shape_name() == lat.shape_name()
