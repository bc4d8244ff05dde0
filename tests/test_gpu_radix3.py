"""Lattices of N = 3 * 2^k points per edge (48, 96, 192, 384): the line
transforms start with an in-register length-12 DFT (a radix-3 step over
length-4 transforms), against the FP64 oracle, whose own transforms of these
lengths are pinned by brute-force DFTs and numpy (test_oracle_kat.py)."""
import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import PSC_VV2, PSC_Y4, PSC_FP32, PSC_FP64, PSC_GW_PROJ_TT
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err, fluct_rel_err

pytestmark = pytest.mark.gpu


def pair(**kw):
    cfg, model, ic = parity_setup(**kw)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    for b in range(cfg.batch):
        lat.set_model(b, model)
        lat.load_ic_k(b, fk, pk, sc[0], sc[1])
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    return cfg, model, ic, lat, orc


@pytest.mark.parametrize("N,nflds,integrator,steps", [
    (48, 2, PSC_VV2, 10), (48, 1, PSC_Y4, 10), (96, 2, PSC_Y4, 10), (192, 2, PSC_VV2, 10), (384, 2, PSC_Y4, 2)])
def test_fp64_step_parity(N, nflds, integrator, steps):
    print(f"case N={N} nflds={nflds} integrator={integrator} steps={steps}", flush=True)
    cfg, model, ic, lat, orc = pair(N=N, nflds=nflds, integrator=integrator)
    lat.step(steps)
    orc.step(steps)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    ef, ep, efl = rel_err(gk, ok), rel_err(gp, op), fluct_rel_err(gk, ok)
    print(f"field {ef:.3e} (fluct {efl:.3e}) momentum {ep:.3e}")
    assert ef <= 1e-10 and ep <= 1e-10 and efl <= 1e-10
    np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10, atol=0)
    lat.close()


def test_fp32_step_error_at_96():
    cfg, model, ic, lat, orc = pair(N=96, integrator=PSC_Y4, precision=PSC_FP32)
    lat.step(10)
    orc.step(10)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    ef, ep = rel_err(gk, ok), rel_err(gp, op)
    print(f"fp32 96: field {ef:.3e} momentum {ep:.3e}")
    assert ef <= 1e-5 and ep <= 1e-5
    lat.close()


def test_energies_spectra_and_fields_at_96():
    cfg, model, ic, lat, orc = pair(N=96, integrator=PSC_Y4)
    lat.step(4)
    orc.step(4)
    np.testing.assert_allclose(lat.energies(0)[:12], orc.energies()[:12], rtol=1e-10, atol=1e-300)
    for fld in (0, 1):
        s, so = lat.spectra(fld, 0), orc.spectra(fld)
        np.testing.assert_array_equal(s[:, 4], so[:, 4])
        np.testing.assert_allclose(s, so, rtol=1e-10, atol=0)
    f, p = lat.fields(0)
    fo, po = orc.fields()
    assert rel_err(f, fo) < 1e-12 and rel_err(p, po) < 1e-12
    lat.close()


def test_device_ic_and_graph_replay_at_96():
    """Device-generated modes against the oracle's; captured steps equal
    eager ones bit for bit."""
    cfg, model, ic = parity_setup(N=96, integrator=PSC_Y4, use_graph=1, omega_star=2e-2, kIR=4.0, g2=2e3)
    ofk, opk, osc = O.generate_ic(cfg, model, ic)
    g = Lattice(cfg)
    g.set_model(0, model)
    g.generate_ic_device(0, ic)
    dfk, dpk = g.get_k()
    e = max(float(np.abs(dfk - ofk).max() / np.abs(ofk).max()), float(np.abs(dpk - opk).max() / np.abs(opk).max()))
    print(f"device IC vs oracle {e:.2e}")
    assert e <= 1e-14
    cfg_e, _, _ = parity_setup(N=96, integrator=PSC_Y4, use_graph=0, omega_star=2e-2, kIR=4.0, g2=2e3)
    eg = Lattice(cfg_e)
    eg.set_model(0, model)
    eg.generate_ic_device(0, ic)
    g.step(5)
    eg.step(5)
    a, b = g.get_k(), eg.get_k()
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    g.close()
    eg.close()


def test_gw_step_and_spectrum_at_48():
    cfg, model, ic = parity_setup(N=48, with_gw=1, integrator=PSC_Y4)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    u, pu = O.generate_gw_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    lat.load_gw_k(0, u, pu)
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    orc.load_gw(u, pu)
    lat.step(6)
    orc.step(6)
    gu, gpu_ = lat.get_gw_k(0)
    ou, opu = orc.get_gw()
    eu, ep = rel_err(gu, ou), rel_err(gpu_, opu)
    print(f"GW 48: u {eu:.3e} pi_u {ep:.3e}")
    assert eu <= 1e-10 and ep <= 1e-10
    s, tot = lat.gw_spectrum(PSC_GW_PROJ_TT, 0)
    so, toto = orc.gw_spectrum(PSC_GW_PROJ_TT)
    np.testing.assert_allclose(s, so, rtol=1e-10, atol=0)
    lat.close()


def test_ensemble_members_at_48():
    cfg, model, ic, lat, orc = pair(N=48, batch=3, integrator=PSC_VV2)
    lat.step(6)
    orc.step(6)
    ok, op = orc.get_k()
    for m in range(3):
        gk, gp = lat.get_k(m)
        assert rel_err(gk, ok) <= 1e-10 and rel_err(gp, op) <= 1e-10
    lat.close()
