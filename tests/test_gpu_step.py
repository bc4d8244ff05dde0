"""GPU parity of the HIP lattice step against the FP64 CPU oracle.

Tolerances: FP64 path <= 1e-10 relative field error after 10 steps, FP32-mixed
path <= 1e-5 relative to the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import PSC_VV2, PSC_Y4, PSC_FP64, PSC_FP32, PSC_EXPANSION_STATIC
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err, fluct_rel_err

pytestmark = pytest.mark.gpu


def _pair(**kw):
    cfg, model, ic = parity_setup(**kw)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    for b in range(cfg.batch):
        lat.set_model(b, model)
        lat.load_ic_k(b, fk, pk, sc[0], sc[1])
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    return cfg, lat, orc, fk, pk, sc


def test_load_roundtrip_fp64():
    cfg, lat, orc, fk, pk, sc = _pair(N=16)
    gk, gp = lat.get_k(0)
    assert rel_err(gk, fk) < 1e-15
    assert rel_err(gp, pk) < 1e-15


def test_initial_add_matches_oracle():
    """d2a/dt2 from the loaded state (initialize.cpp:797) vs the oracle."""
    cfg, lat, orc, fk, pk, sc = _pair(N=32)
    s = lat.scalars(0)
    print("gpu add", s[3], "oracle", sc[4])
    assert abs(s[3] - sc[4]) <= 1e-12 * abs(sc[4])


@pytest.mark.parametrize("N", [16, 32, 64])
def test_real_fields_match_oracle(N):
    cfg, lat, orc, fk, pk, sc = _pair(N=N)
    f, p = lat.fields(0)
    fo, po = orc.fields()
    print(N, rel_err(f, fo), rel_err(p, po))
    assert rel_err(f, fo) < 1e-13
    assert rel_err(p, po) < 1e-13


@pytest.mark.parametrize("integrator", [PSC_VV2, PSC_Y4])
@pytest.mark.parametrize("nflds", [1, 2])
@pytest.mark.parametrize("N", [16, 32, 64])
def test_fp64_step_parity(N, nflds, integrator):
    """10 steps of the FP64 GPU path vs the oracle: <= 1e-10 relative."""
    cfg, lat, orc, fk, pk, sc = _pair(N=N, nflds=nflds, integrator=integrator)
    lat.step(10)
    orc.step(10)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    ef, ep = rel_err(gk, ok), rel_err(gp, op)
    eff = fluct_rel_err(gk, ok)
    s, so = lat.scalars(0), orc.scalars()
    print(f"N={N} nflds={nflds} int={integrator} field {ef:.3e} (fluct {eff:.3e}) mom {ep:.3e} "
          f"scalars gpu {s} oracle {so}")
    assert ef <= 1e-10 and ep <= 1e-10 and eff <= 1e-10
    np.testing.assert_allclose(s, so, rtol=1e-10, atol=0)


def test_fp64_step_parity_split_calls():
    """psc_step(3) + psc_step(7) equals 10 oracle steps (cached scalars carry over)."""
    cfg, lat, orc, fk, pk, sc = _pair(N=32, integrator=PSC_Y4)
    lat.step(3)
    lat.step(7)
    orc.step(10)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    assert rel_err(gk, ok) <= 1e-10 and rel_err(gp, op) <= 1e-10
    np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10)


@pytest.mark.parametrize("integrator", [PSC_VV2, PSC_Y4])
def test_fp32_step_error(integrator):
    """FP32-mixed path vs the FP64 oracle over 10 steps: <= 1e-5 relative."""
    cfg, lat, orc, fk, pk, sc = _pair(N=32, integrator=integrator, precision=PSC_FP32)
    lat.step(10)
    orc.step(10)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    ef, ep, eff = rel_err(gk, ok), rel_err(gp, op), fluct_rel_err(gk, ok)
    s, so = lat.scalars(0), orc.scalars()
    print(f"fp32 int={integrator}: field {ef:.3e} fluct {eff:.3e} mom {ep:.3e} scalars {s} vs {so}")
    assert ef <= 1e-5 and ep <= 1e-5
    np.testing.assert_allclose(s[1:3], so[1:3], rtol=1e-6)


def test_energies_match_oracle():
    cfg, lat, orc, fk, pk, sc = _pair(N=32, integrator=PSC_Y4)
    lat.step(5)
    orc.step(5)
    e, eo = lat.energies(0), orc.energies()
    print("gpu", e)
    print("orc", eo)
    np.testing.assert_allclose(e[:12], eo[:12], rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(e[12:], eo[12:], rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("fld", [0, 1])
def test_spectra_match_oracle(fld):
    cfg, lat, orc, fk, pk, sc = _pair(N=32)
    lat.step(4)
    orc.step(4)
    s, so = lat.spectra(fld, 0), orc.spectra(fld)
    np.testing.assert_array_equal(s[:, 4], so[:, 4])
    np.testing.assert_allclose(s, so, rtol=1e-10, atol=0)


def test_batch_members_match_single_runs():
    """Three members with different couplings and seeds in one context equal
    three single-member runs bit for bit."""
    from pspectcosmo_amd.abi import make_model, make_ic
    cfg, model, ic = parity_setup(N=16, batch=3, integrator=PSC_Y4)
    lat = Lattice(cfg)
    ref = []
    for b in range(3):
        m = make_model(1.0, [1e4, 1e5, 2.5e5][b])
        i = make_ic(f0=(1.0 + 0.1 * b, 0.0), fd0=(-0.7, 0.0), seed=b + 1)
        fk, pk, sc = O.generate_ic(cfg, m, i)
        lat.set_model(b, m)
        lat.load_ic_k(b, fk, pk, sc[0], sc[1])
        c1, _, _ = parity_setup(N=16, batch=1, integrator=PSC_Y4)
        one = Lattice(c1)
        one.set_model(0, m)
        one.load_ic_k(0, fk, pk, sc[0], sc[1])
        one.step(6)
        ref.append((one.get_k(0), one.scalars(0)))
    lat.step(6)
    for b in range(3):
        gk, gp = lat.get_k(b)
        (rk, rp), rs = ref[b]
        np.testing.assert_array_equal(gk, rk)
        np.testing.assert_array_equal(gp, rp)
        np.testing.assert_array_equal(lat.scalars(b), rs)


def test_graph_replay_matches_eager():
    """hipGraph-captured steps equal eager launches bit for bit, replay after replay."""
    cfg, lat, orc, fk, pk, sc = _pair(N=32, integrator=PSC_Y4)
    cfg_g, _, _ = parity_setup(N=32, integrator=PSC_Y4, use_graph=1)
    g = Lattice(cfg_g)
    from pspectcosmo_amd.abi import make_model
    g.set_model(0, make_model(1.0, 2.5e5))
    g.load_ic_k(0, fk, pk, sc[0], sc[1])
    for rep in range(3):
        lat.step(4)
        g.step(4)
        g.sync()
        a, b = lat.get_k(0), g.get_k(0)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(lat.scalars(0), g.scalars(0))


def test_static_free_field_exact_solution():
    """Minkowski free field: every mode follows the exact discrete VV map."""
    N, kIR, m2, dt, nsteps = 16, 1.0, 4.0, 0.05, 40
    cfg, model, ic = parity_setup(N=N, nflds=1, kIR=kIR, dt=dt, g2=0.0, m2=m2,
                                  expansion=PSC_EXPANSION_STATIC, omega_star=1e-2,
                                  fd0=(0.3, 0.0))
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, 1.0, 0.0)
    lat.step(nsteps)
    gk, gp = lat.get_k(0)
    i = np.arange(N)
    n = np.where(i <= N // 2, i, i - N)
    nz = np.arange(N // 2 + 1)
    n2 = n[:, None, None] ** 2 + n[None, :, None] ** 2 + nz[None, None, :] ** 2
    w2 = kIR ** 2 * n2 + m2
    A = 1 - w2 * dt * dt / 2
    Bm = dt
    Cm = -w2 * dt * (1 - w2 * dt * dt / 4)
    f, p = fk[0].copy(), pk[0].copy()
    for _ in range(nsteps):
        f, p = A * f + Bm * p, Cm * f + A * p
    print(rel_err(gk[0], f), rel_err(gp[0], p))
    assert rel_err(gk[0], f) < 1e-12 and rel_err(gp[0], p) < 1e-12
