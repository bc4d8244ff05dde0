"""The model.h boundary on the GPU: a second potential, V = lambda phi^4 / 4
+ g2 phi^2 chi^2 / 2 (models/lambda_phi4.hpp, built into
libpsc_lambda_phi4.so with no other source change), against the oracle built
with the same header: step, energies, spectra and adjoint parity, as the
preheat model's tests check them.  Also the Friedmann-drift configuration of
test_oracle_dynamics.py (gradient energy >= 10 % of rho, a doubling) on the
GPU for both models."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import (PSC_VV2, PSC_Y4, PSC_FP64, PSC_FP32, PSC_COUPLING_G2, PSC_COUPLING_LAMBDA,
                                 make_model)
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err, fluct_rel_err

pytestmark = pytest.mark.gpu

NAME = "lambda_phi4"


def quartic():
    return make_model(m2=0.0, g2=120.0, lam=1.0)


def _pair(name=NAME, model=None, **kw):
    cfg, _, ic = parity_setup(**kw)
    model = model or quartic()
    fk, pk, sc = O.generate_ic(cfg, model, ic, model_name=name)
    lat = Lattice(cfg, model_name=name)
    assert lat._L.psc_model_name().decode() == name
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    orc = O.Oracle(cfg, model, model_name=name)
    orc.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    return cfg, lat, orc, fk, pk, sc


@pytest.mark.parametrize("integrator", [PSC_VV2, PSC_Y4])
@pytest.mark.parametrize("nflds", [1, 2])
@pytest.mark.parametrize("N", [16, 32])
def test_fp64_step_parity(N, nflds, integrator):
    cfg, lat, orc, fk, pk, sc = _pair(N=N, nflds=nflds, integrator=integrator)
    lat.step(10)
    orc.step(10)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    ef, ep, eff = rel_err(gk, ok), rel_err(gp, op), fluct_rel_err(gk, ok)
    print(f"{NAME} N={N} nflds={nflds} int={integrator}: field {ef:.3e} fluct {eff:.3e} mom {ep:.3e}", flush=True)
    assert ef <= 1e-10 and ep <= 1e-10 and eff <= 1e-10
    np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10, atol=0)


def test_256_fp64_and_fp32_parity():
    """256^3, two Yoshida-4 steps: FP64 <= 1e-10, FP32-mixed <= 1e-5 against
    the oracle."""
    cfg, lat, orc, fk, pk, sc = _pair(N=256, integrator=PSC_Y4)
    lat.step(2)
    orc.step(2)
    ok, op = orc.get_k()
    gk, gp = lat.get_k(0)
    print("256 fp64", rel_err(gk, ok), rel_err(gp, op), flush=True)
    assert rel_err(gk, ok) <= 1e-10 and rel_err(gp, op) <= 1e-10
    lat.close()
    cfg32, _, ic = parity_setup(N=256, integrator=PSC_Y4, precision=PSC_FP32)
    lat32 = Lattice(cfg32, model_name=NAME)
    lat32.set_model(0, quartic())
    lat32.load_ic_k(0, fk, pk, sc[0], sc[1])
    lat32.step(2)
    gk, gp = lat32.get_k(0)
    print("256 fp32", rel_err(gk, ok), fluct_rel_err(gk, ok), rel_err(gp, op), flush=True)
    assert rel_err(gk, ok) <= 1e-5 and rel_err(gp, op) <= 1e-5 and fluct_rel_err(gk, ok) <= 1e-5


def test_energies_terms_and_spectra():
    cfg, lat, orc, fk, pk, sc = _pair(N=32, integrator=PSC_Y4)
    lat.step(5)
    orc.step(5)
    e, eo = lat.energies(0), orc.energies()
    np.testing.assert_allclose(e[:12], eo[:12], rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(lat.potential_terms(0), eo[4:6], rtol=1e-10)
    for fld in (0, 1):
        s, so = lat.spectra(fld, 0), orc.spectra(fld)
        np.testing.assert_array_equal(s[:, 4], so[:, 4])
        np.testing.assert_allclose(s, so, rtol=1e-10, atol=0)


@pytest.mark.parametrize("coupling", [PSC_COUPLING_LAMBDA, PSC_COUPLING_G2])
def test_fp64_adjoint_matches_oracle(coupling):
    phi0, dphi0 = 1.0, -0.7
    cfg, _, ic = parity_setup(N=16, nflds=2, integrator=PSC_Y4, kIR=4.0, dt=0.01, omega_star=2e-2,
                              f0=(phi0, 0.0), fd0=(dphi0, 0.0), yoshida_exact=1)
    model = make_model(m2=0.0, g2=2.0e3, lam=20.0)
    fk, pk, sc = O.generate_ic(cfg, model, ic, model_name=NAME)
    lat = Lattice(cfg, model_name=NAME)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    g = lat.adjoint(7, 3, phi0, dphi0, coupling=coupling)
    J, dc, dp, dchi = O.adjoint_chi2(cfg, model, fk, pk, 1.0, phi0, dphi0, 0.0, 7, coupling=coupling,
                                     model_name=NAME)
    print("adjoint", coupling, g["J"], J, g["dg2"], dc, g["dphi0"], dp, flush=True)
    assert abs(g["J"] - J) <= 1e-12 * abs(J)
    assert abs(g["dg2"] - dc) <= 1e-9 * abs(dc)
    assert abs(g["dphi0"] - dp) <= 1e-9 * abs(dp)
    assert rel_err(g["dchi"], dchi) <= 1e-9
    with pytest.raises(RuntimeError):   # the model has no m2 derivative
        lat.adjoint(2, 1, phi0, dphi0, coupling=1)
    lat.close()


@pytest.mark.parametrize("name,model", [("preheat", make_model(m2=1.0, g2=50.0)),
                                        ("lambda_phi4", make_model(m2=0.0, g2=50.0, lam=2.0))])
def test_friedmann_configuration_parity(name, model):
    """The configuration whose constraint drift pins the oracle's expanding
    dynamics (kIR = 1, omegaStar = 0.1: gradient energy ~12 % of rho; da/dt
    from the full rho): FP64 GPU vs oracle after 40 VV steps, fields,
    scalars and the constraint violation."""
    cfg, _, ic = parity_setup(N=16, nflds=2, integrator=PSC_VV2, dt=0.01, kIR=1.0, omega_star=0.1,
                              f0=(1.0, 0.3), fd0=(-0.7, 0.2))
    fk, pk, sc = O.generate_ic(cfg, model, ic, model_name=name)
    orc = O.Oracle(cfg, model, model_name=name)
    orc.load(fk, pk, 0.0, sc[0], sc[1], 0.0)
    e0 = orc.energies()
    assert e0[1] + e0[3] >= 0.1 * e0[6]
    ad0 = sc[0] * math.sqrt(e0[6] / 3.0)
    orc.load(fk, pk, 0.0, sc[0], ad0, 0.0)
    add0 = orc.kernel_a()
    orc.load(fk, pk, 0.0, sc[0], ad0, add0)
    lat = Lattice(cfg, model_name=name)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], ad0)
    lat.step(40)
    orc.step(40)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    e, eo = lat.energies(0), orc.energies()
    print(name, rel_err(gk, ok), rel_err(gp, op), "constraint", e[7], eo[7], "a", e[8], flush=True)
    assert rel_err(gk, ok) <= 1e-10 and rel_err(gp, op) <= 1e-10
    np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10)
    assert abs(e[7] - eo[7]) <= 1e-6 * eo[7] + 1e-15
