"""GPU parity of the gravitational-wave evolution and spectra against the FP64
oracle (same seeded fields and GW initial modes)."""
import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import (PSC_VV2, PSC_Y4, PSC_FP64, PSC_FP32, PSC_GW_PROJ_REFERENCE,
                                 PSC_GW_PROJ_TT)
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err

pytestmark = pytest.mark.gpu


def _pair(**kw):
    cfg, model, ic = parity_setup(with_gw=1, **kw)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    u, pu = O.generate_gw_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    lat.load_gw_k(0, u, pu)
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    orc.load_gw(u, pu)
    return cfg, lat, orc, u, pu


def test_gw_load_roundtrip():
    cfg, lat, orc, u, pu = _pair(N=16)
    gu, gpu_ = lat.get_gw_k(0)
    assert rel_err(gu, u) < 1e-15 and rel_err(gpu_, pu) < 1e-15


@pytest.mark.parametrize("integrator", [PSC_VV2, PSC_Y4])
@pytest.mark.parametrize("N", [16, 32])
def test_gw_step_parity_fp64(N, integrator):
    cfg, lat, orc, u, pu = _pair(N=N, integrator=integrator)
    lat.step(10)
    orc.step(10)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    gu, gpu_ = lat.get_gw_k(0)
    ou, opu = orc.get_gw()
    errs = [rel_err(gk, ok), rel_err(gp, op), rel_err(gu, ou), rel_err(gpu_, opu)]
    print(f"N={N} int={integrator} field/mom/u/pi_u errors {errs}")
    assert max(errs) <= 1e-10
    np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10)


@pytest.mark.parametrize("projection", [PSC_GW_PROJ_REFERENCE, PSC_GW_PROJ_TT])
def test_gw_spectrum_parity(projection):
    cfg, lat, orc, u, pu = _pair(N=16, integrator=PSC_Y4)
    lat.step(4)
    orc.step(4)
    s, tot = lat.gw_spectrum(projection, 0)
    so, toto = orc.gw_spectrum(projection)
    np.testing.assert_array_equal(s[:, 4], so[:, 4])
    np.testing.assert_allclose(s, so, rtol=1e-10, atol=0)
    np.testing.assert_allclose(tot, toto, rtol=1e-10)
    kc, om = lat.gw_omega_log(projection, 64, 0)
    kco, omo = orc.gw_omega_log(projection, 64)
    np.testing.assert_allclose(kc, kco, rtol=1e-14)
    np.testing.assert_allclose(om, omo, rtol=1e-10, atol=1e-300)


def test_gw_step_fp32_error():
    cfg, lat, orc, u, pu = _pair(N=32, integrator=PSC_Y4, precision=PSC_FP32)
    lat.step(10)
    orc.step(10)
    gu, gpu_ = lat.get_gw_k(0)
    ou, opu = orc.get_gw()
    gk, _ = lat.get_k(0)
    ok, _ = orc.get_k()
    errs = [rel_err(gk, ok), rel_err(gu, ou), rel_err(gpu_, opu)]
    print("fp32 gw errors", errs)
    assert max(errs) <= 1e-5


def test_gw_graph_replay_matches_eager():
    cfg, lat, orc, u, pu = _pair(N=16, integrator=PSC_Y4)
    cfg_g, model, ic = parity_setup(N=16, integrator=PSC_Y4, with_gw=1, use_graph=1)
    fk, pk, sc = O.generate_ic(cfg_g, model, ic)
    g = Lattice(cfg_g)
    g.set_model(0, model)
    g.load_ic_k(0, fk, pk, sc[0], sc[1])
    g.load_gw_k(0, u, pu)
    for _ in range(2):
        lat.step(3)
        g.step(3)
        a, b = lat.get_gw_k(0), g.get_gw_k(0)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


def test_gw_512_shapes_agree():
    """At 512^3 the default shape runs the y passes, the GW y inverse among
    them, on a second plan (radix 16/16/2, 32 columns per block); every 512
    shape gives the same change of the GW and field spectra over four
    Yoshida-4 steps, FP32 (the change, not the spectra themselves, which the
    initial vacuum modes dominate)."""
    cfg, model, ic = parity_setup(N=512, with_gw=1, integrator=PSC_Y4, precision=PSC_FP32)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    shapes = lat.shapes()
    print("512 GW shapes", shapes, flush=True)
    assert "_y" in shapes[0]
    ref = None
    for name in shapes:
        lat.set_shape(name)
        lat.generate_ic_device(0, ic)
        lat.generate_gw_ic_device(0, ic)
        s0, _ = lat.gw_spectrum(PSC_GW_PROJ_TT, 0)
        f0 = lat.spectra(0, 0)
        print(f"{name}: initial scalars {lat.scalars(0)}, field spectrum sum {f0.sum():.12e}, "
              f"GW spectrum sum {s0.sum():.12e}", flush=True)
        lat.step(4)
        s1, _ = lat.gw_spectrum(PSC_GW_PROJ_TT, 0)
        f1 = lat.spectra(0, 0)
        ds, df = s1 - s0, f1 - f0
        print(f"{name}: scalars after the steps {lat.scalars(0)}", flush=True)
        if ref is None:
            ref = (ds, df)
            print(f"{name}: GW spectrum change {np.abs(ds).max() / np.abs(s0).max():.2e} of its size, "
                  f"field spectrum {np.abs(df).max() / np.abs(f0).max():.2e}", flush=True)
            assert np.all(np.isfinite(s1)) and np.abs(ds).max() > 1e-9 * np.abs(s0).max()
            continue
        es = float(np.abs(ds - ref[0]).max() / np.abs(ref[0]).max())
        ef = float(np.abs(df - ref[1]).max() / np.abs(ref[1]).max())
        print(f"{name}: GW change {es:.2e} field change {ef:.2e}", flush=True)
        assert es <= 1e-3 and ef <= 1e-3, name
    lat.close()


@pytest.mark.parametrize("precision,tol", [(PSC_FP64, 1e-12), (PSC_FP32, 2e-6)])
def test_free_gw_exact_at_512(precision, tol):
    """The 512^3 GW instantiations against the exact discrete solution: with
    phi = chi = 0 the source d_i f d_j f vanishes, and on a static background
    each component of each mode follows the velocity-Verlet map of
    u'' = -kIR^2 n^2 u (lapl, evolution.cpp:343-372) exactly:
    u' = A u + dt pi, pi' = C u + A pi with A = 1 - w^2 dt^2 / 2,
    C = -w^2 dt (1 - w^2 dt^2 / 4).  Random modes in all six components
    (the kz = 0 and N/2 planes left empty so that no Hermitian pairing is
    needed), two steps at k_max dt = 1.1."""
    from pspectcosmo_amd.abi import PSC_EXPANSION_STATIC
    N, kIR, dt, steps = 512, 0.5, 0.005, 2
    cfg, model, ic = parity_setup(N=N, with_gw=1, integrator=PSC_VV2, precision=precision, kIR=kIR, dt=dt,
                                  expansion=PSC_EXPANSION_STATIC)
    print(f"case free GW 512^3 precision {precision}", flush=True)
    rng = np.random.default_rng(3)
    shape = (6, N, N, N // 2 + 1)
    u = np.empty(shape, np.complex128)
    pu = np.empty(shape, np.complex128)
    for c in range(6):
        for arr in (u, pu):
            arr[c].real = rng.standard_normal(shape[1:], dtype=np.float32)
            arr[c].imag = rng.standard_normal(shape[1:], dtype=np.float32)
            arr[c][:, :, 0] = 0
            arr[c][:, :, N // 2] = 0
    fk = np.zeros((2, N, N, N // 2 + 1), np.complex128)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, fk, 1.0, 0.0)
    del fk
    lat.load_gw_k(0, u, pu)
    print("loaded", flush=True)
    lat.step(steps)
    gu, gp = lat.get_gw_k(0)
    lat.close()
    i = np.arange(N)
    n = np.where(i <= N // 2, i, i - N).astype(np.float64)
    nz = np.arange(N // 2 + 1, dtype=np.float64)
    w2 = kIR ** 2 * (n[:, None, None] ** 2 + n[None, :, None] ** 2 + nz[None, None, :] ** 2)
    A = 1 - w2 * dt * dt / 2
    C = -w2 * dt * (1 - w2 * dt * dt / 4)
    num_u = num_p = den_u = den_p = 0.0
    for c in range(6):
        f, p = u[c], pu[c]
        for _ in range(steps):
            f, p = A * f + dt * p, C * f + A * p
        num_u += np.linalg.norm(gu[c] - f) ** 2
        num_p += np.linalg.norm(gp[c] - p) ** 2
        den_u += np.linalg.norm(f) ** 2
        den_p += np.linalg.norm(p) ** 2
    eu, ep = float(np.sqrt(num_u / den_u)), float(np.sqrt(num_p / den_p))
    print(f"free GW 512^3: u {eu:.3e} pi_u {ep:.3e}")
    assert eu < tol and ep < tol
