"""The dealiased stepper (cfg.pad = 2: evolve_VV_k_pad, evolution.cpp:1271-1759,
N_pad = 2N and 3N/2) on the GPU against the oracle's restatement: half spectra of
f and pi after 10 steps including the anti-Hermitian parts the reference's
extraction leaves in the kz = 0, N/2 planes, the background scalars, K and G
and the binned spectra of both fields.  FP64 <= 1e-10, FP32 <= 1e-5 (relative L2 over the
spectrum).  Also: graph-captured steps equal eager ones, the linear force
reproduces the unpadded stepper, and unsupported combinations are rejected."""
import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import (PSC_FP32, PSC_FP64, PSC_VV2, PSC_Y4, PSC_EXPANSION_STATIC, make_config,
                                 make_model)
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err, fluct_rel_err

pytestmark = pytest.mark.gpu


def pair(N, integrator, precision, steps, use_graph=0, g2=2e3, n_pad=0, **kw):
    cfg, model, ic = parity_setup(N=N, integrator=integrator, precision=precision, omega_star=2e-2, kIR=4.0,
                                  g2=g2, pad=2, n_pad=n_pad, use_graph=use_graph, **kw)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, sc[0], sc[1], 0.0)
    add = o.kernel_a()
    o.load(fk, pk, 0.0, sc[0], sc[1], add)
    np.testing.assert_allclose(lat.scalars(), o.scalars(), rtol=1e-12 if precision == PSC_FP64 else 1e-6)
    lat.step(steps)
    o.step(steps)
    return lat, o


def ghost_size(fk):
    """Largest anti-Hermitian part in the kz = 0 plane of the half spectra."""
    N = fk.shape[1]
    p = fk[..., 0]
    m = np.conj(p[:, (-np.arange(N)) % N][:, :, (-np.arange(N)) % N])
    return float(np.abs(p - m).max() / np.abs(p).max())


@pytest.mark.parametrize("N,integrator,n_pad,steps", [
    (16, PSC_VV2, 0, 10), (32, PSC_Y4, 0, 10), (64, PSC_VV2, 0, 10),
    (32, PSC_Y4, 48, 10), (64, PSC_VV2, 96, 10), (128, PSC_Y4, 192, 3),   # N_pad = 3N/2
    (48, PSC_Y4, 96, 6)])                                                 # N_pad = 2N on a 3 * 2^k grid
def test_fp64_padded_matches_oracle(N, integrator, n_pad, steps):
    print("case padded fp64", N, "integrator", integrator, "N_pad", n_pad or 2 * N, flush=True)
    lat, o = pair(N, integrator, PSC_FP64, steps, n_pad=n_pad)
    (gf, gp), (of, op) = lat.get_k(), o.get_k()
    ef, ep, efl = rel_err(gf, of), rel_err(gp, op), fluct_rel_err(gf, of)
    print(f"field {ef:.2e} (fluctuations {efl:.2e}) momentum {ep:.2e}; ghost of the oracle {ghost_size(of):.2e}")
    assert ghost_size(of) > 1e-8      # the asymmetric extraction is exercised
    assert ef <= 1e-10 and ep <= 1e-10 and efl <= 1e-10
    np.testing.assert_allclose(lat.scalars(), o.scalars(), rtol=1e-12)
    np.testing.assert_allclose(lat.energies()[:4], o.energies()[:4], rtol=1e-10)
    for f in range(2):   # the binned spectra of the half spectra, ghosts included
        np.testing.assert_allclose(lat.spectra(f), o.spectra(f), rtol=1e-9, atol=0)
    lat.close()


@pytest.mark.parametrize("N,n_pad", [(32, 0), (64, 96)])
def test_fp32_padded_matches_oracle(N, n_pad):
    print("case padded fp32 Y4", N, "N_pad", n_pad or 2 * N, flush=True)
    lat, o = pair(N, PSC_Y4, PSC_FP32, 10, n_pad=n_pad)
    (gf, gp), (of, op) = lat.get_k(), o.get_k()
    ef, ep, efl = rel_err(gf, of), rel_err(gp, op), fluct_rel_err(gf, of)
    print(f"field {ef:.2e} (fluctuations {efl:.2e}) momentum {ep:.2e}")
    assert ef <= 1e-5 and ep <= 1e-5 and efl <= 1e-5
    np.testing.assert_allclose(lat.scalars(), o.scalars(), rtol=1e-6)
    lat.close()


@pytest.mark.parametrize("n_pad", [0, 48])
def test_graph_equals_eager(n_pad):
    a, _ = pair(32, PSC_Y4, PSC_FP64, 6, use_graph=0, n_pad=n_pad)
    b, _ = pair(32, PSC_Y4, PSC_FP64, 6, use_graph=1, n_pad=n_pad)
    (af, ap), (bf, bp) = a.get_k(), b.get_k()
    assert np.array_equal(af, bf) and np.array_equal(ap, bp)
    assert np.array_equal(a.scalars(), b.scalars())
    a.close()
    b.close()


@pytest.mark.parametrize("N,precision,n_pad,tol,tol_fl", [
    (32, PSC_FP64, 0, 1e-13, 1e-12), (32, PSC_FP64, 48, 1e-13, 1e-12), (256, PSC_FP32, 384, 2e-6, 1e-5),
    (512, PSC_FP32, 768, 2e-6, 1e-5), (384, PSC_FP64, 768, 1e-13, 1e-12)])
def test_linear_force_matches_unpadded_static(N, precision, n_pad, tol, tol_fl):
    """A linear V' has no products to alias: the padded stepper (here N_pad =
    2N, and 3N/2 up to the 384-point plans) reproduces the unpadded one."""
    print("case linear force padded", N, "N_pad", n_pad or 2 * N, flush=True)
    kw = dict(N=N, integrator=PSC_Y4, precision=precision, omega_star=2e-2, kIR=4.0, g2=0.0, m2=2.0,
              expansion=PSC_EXPANSION_STATIC)
    out = []
    for pad in (0, 2):
        cfg, model, ic = parity_setup(pad=pad, n_pad=n_pad if pad else 0, **kw)
        fk, pk, sc = O.generate_ic(cfg, model, ic)
        lat = Lattice(cfg)
        lat.set_model(0, model)
        lat.load_ic_k(0, fk, pk, 1.0, 0.0)
        lat.step(5 if N <= 64 else 2)
        out.append(lat.get_k())
        lat.close()
    # the fluctuations (k = 0 removed) as well: the homogeneous mode is exact
    # in both and would hide an error in the transforms
    e = max(rel_err(out[1][0], out[0][0]), rel_err(out[1][1], out[0][1]))
    efl = max(fluct_rel_err(out[1][0], out[0][0]), fluct_rel_err(out[1][1], out[0][1]))
    print(f"padded against unpadded {e:.2e} (fluctuations {efl:.2e})")
    assert e < tol and efl < tol_fl


def test_nonlinear_768_padded_step_is_time_reversible():
    """The 512^3 lattice dealiased on 768^3 (N_pad = 3N/2), FP64, static
    background: two Yoshida-4 steps forward, then two with -dt from the
    result in a second context, return the state to rounding; the forward
    steps changed it (nonlinear V')."""
    print("case padded 512 on 768, fp64, time reversal", flush=True)
    # a vacuum amplitude small enough at 512^3 that two steps stay gentle (with
    # 2e-2 the UV modes grow 20 % per two steps and amplify the rounding)
    kw = dict(N=512, integrator=PSC_Y4, precision=PSC_FP64, omega_star=1e-5, kIR=4.0, g2=2e3, pad=2, n_pad=768,
              expansion=PSC_EXPANSION_STATIC)
    cfg, model, ic = parity_setup(**kw)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.generate_ic_device(0, ic)
    f0, p0 = lat.get_k()
    lat.step(2)
    f1, p1 = lat.get_k()
    lat.close()
    cfg_b, _, _ = parity_setup(dt=-cfg.dt, **kw)
    back = Lattice(cfg_b)
    back.set_model(0, model)
    back.load_ic_k(0, f1, p1, 1.0, 0.0)
    back.step(2)
    f2, p2 = back.get_k()
    back.close()
    moved, ret, retp = rel_err(f1, f0), rel_err(f2, f0), rel_err(p2, p0)
    print(f"forward change {moved:.2e}; after the return: field {ret:.2e} momentum {retp:.2e}")
    assert moved > 1e-8 and ret < 1e-12 and retp < 1e-12


@pytest.mark.parametrize("n_pad", [0, 48])
def test_get_load_round_trip_keeps_the_plane_ghosts(n_pad):
    """After steps the half spectra of the dealiased stepper carry anti-Hermitian
    parts in the kz = 0, N/2 planes; psc_get_k -> psc_load_ic_k in a fresh
    context reads them back and the two contexts step on alike."""
    print("case padded get/load round trip, N_pad", n_pad or 64, flush=True)
    lat, _ = pair(32, PSC_Y4, PSC_FP64, 10, n_pad=n_pad, expansion=PSC_EXPANSION_STATIC)
    f1, p1 = lat.get_k()
    assert ghost_size(f1) > 1e-8
    cfg, model, _ = parity_setup(N=32, integrator=PSC_Y4, precision=PSC_FP64, omega_star=2e-2, kIR=4.0, g2=2e3,
                                 pad=2, n_pad=n_pad, expansion=PSC_EXPANSION_STATIC)
    other = Lattice(cfg)
    other.set_model(0, model)
    other.load_ic_k(0, f1, p1, 1.0, 0.0)
    g1, q1 = other.get_k()
    assert rel_err(g1, f1) < 1e-15 and rel_err(q1, p1) < 1e-15
    lat.step(3)
    other.step(3)
    (af, ap), (bf, bp) = lat.get_k(), other.get_k()
    e = max(rel_err(bf, af), rel_err(bp, ap))
    print(f"after 3 more steps: {e:.2e}")
    assert e < 1e-13
    lat.close()
    other.close()


def test_rejections():
    for kw in (dict(batch=2), dict(with_gw=1), dict(N=512), dict(n_pad=32), dict(n_pad=40), dict(n_pad=80),
               dict(N=512, n_pad=1024), dict(N=384, n_pad=576), dict(pad=0, n_pad=64)):
        args = dict(N=32, pad=2)
        args.update(kw)
        cfg = make_config(args.pop("N"), **args)
        with pytest.raises(Exception):
            Lattice(cfg)
    with pytest.raises(Exception):
        Lattice(make_config(32, pad=1))


def test_padded_shapes_agree_and_autotune_restores_the_state():
    cfg, model, ic = parity_setup(N=64, integrator=PSC_Y4, precision=PSC_FP64, omega_star=2e-2, kIR=4.0,
                                  g2=2e3, pad=2)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    shapes = lat.shapes()
    print("padded shapes", shapes, flush=True)
    assert len(shapes) >= 2 and lat.shape_name() == shapes[0]
    ref = None
    for name in shapes:
        lat.set_shape(name)
        lat.load_ic_k(0, fk, pk, sc[0], sc[1])
        lat.step(4)
        got = lat.get_k()
        if ref is None:
            ref = got
        else:
            assert rel_err(got[0], ref[0]) < 1e-12 and rel_err(got[1], ref[1]) < 1e-12, name
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    lat.autotune(1)
    assert lat.shape_name() in shapes
    lat.step(4)
    got = lat.get_k()
    assert rel_err(got[0], ref[0]) < 1e-12 and rel_err(got[1], ref[1]) < 1e-12
    lat.close()


def test_padded_256_shapes_agree():
    """Every 512-point plan of the dealiased 256^3 stepper, the one with its z
    row passes on a second plan included, against the default, FP32."""
    cfg, model, ic = parity_setup(N=256, integrator=PSC_Y4, precision=PSC_FP32, pad=2)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    shapes = lat.shapes()
    print("padded 256 shapes", shapes, flush=True)
    assert any("_z" in s for s in shapes) and lat.shape_name() == shapes[0]
    ref = None
    for name in shapes:
        lat.set_shape(name)
        lat.load_ic_k(0, fk, pk, sc[0], sc[1])
        lat.step(2)
        got = lat.get_k()
        if ref is None:
            ref = got
            continue
        e = max(rel_err(got[0], ref[0]), rel_err(got[1], ref[1]))
        print(f"{name}: {e:.3e}")
        assert e <= 2e-6, name
    lat.close()


def test_reload_after_steps_matches_a_fresh_context():
    """Loading new data clears the plane ghosts left by earlier steps before
    d2a/dt2 is formed from K, G and the padded V."""
    cfg, model, ic = parity_setup(N=32, integrator=PSC_VV2, precision=PSC_FP64, omega_star=2e-2, kIR=4.0,
                                  g2=2e3, pad=2)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    fresh = lat.scalars()
    lat.step(3)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    np.testing.assert_array_equal(lat.scalars(), fresh)
    lat.close()
