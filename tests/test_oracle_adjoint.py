"""Pin the oracle's FP64 discrete adjoint against central finite differences
of the oracle's half-spectrum forward restatement (a separate implementation
of the same step): d<chi^2>/dg2, d<chi^2>/dphi0 and d<chi^2>/dchi(x).

Central differences converge to the adjoint as eps^2 (relative gap 4e-5,
4e-7, 9e-9 at eps = 1e-3, 1e-4, 1e-5 relative for g^2), so the pin uses
eps = 1e-5 relative and a 1e-6 tolerance."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import (PSC_VV2, PSC_Y4, PSC_EXPANSION_STATIC, PSC_EXPANSION_SELF, make_model,
                                 PSC_COUPLING_G2, PSC_COUPLING_LAMBDA)
from tests._util import parity_setup

PHI0, DPHI0 = 1.0, -0.7


def forward_J(cfg, model, fk, pk, phi0, steps, dchi_site=None, eps_chi=0.0):
    """<chi^2> after `steps` steps of the half-spectrum oracle, with the
    initial state (fk, pk), phi0 setting the DC mode and H(phi0)."""
    N = cfg.N
    fk = fk.copy()
    fk[0, 0, 0, 0] = phi0 * N ** 3
    if dchi_site is not None:
        x = np.zeros((N, N, N))
        x[dchi_site] = eps_chi
        fk[1] += O.fft_r2c_3d(x)
    V0 = 0.5 * model.m2 * phi0 ** 2
    H = math.sqrt((DPHI0 ** 2 / 2 + V0) / 3.0)
    stat = cfg.expansion == PSC_EXPANSION_STATIC
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, 1.0, 0.0 if stat else H, 0.0)
    add0 = o.kernel_a()
    o.load(fk, pk, 0.0, 1.0, 0.0 if stat else H, add0)
    o.step(steps)
    f, _ = o.fields()
    return float((f[1] ** 2).mean())


def setup(integrator, expansion, N=8, g2=2.0e3):
    cfg, model, ic = parity_setup(N=N, nflds=2, integrator=integrator, expansion=expansion,
                                  kIR=4.0, dt=0.01, g2=g2, omega_star=2e-2, f0=(PHI0, 0.0),
                                  fd0=(DPHI0, 0.0), yoshida_exact=1)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    return cfg, model, fk, pk


@pytest.mark.parametrize("integrator,expansion", [(PSC_VV2, PSC_EXPANSION_SELF),
                                                  (PSC_Y4, PSC_EXPANSION_SELF),
                                                  (PSC_VV2, PSC_EXPANSION_STATIC)])
def test_adjoint_matches_central_differences(integrator, expansion):
    steps = 6
    cfg, model, fk, pk = setup(integrator, expansion)
    J, dg2, dphi0, dchi = O.adjoint_chi2(cfg, model, fk, pk, 1.0, PHI0, DPHI0, 0.0, steps)
    J0 = forward_J(cfg, model, fk, pk, PHI0, steps)
    assert abs(J - J0) <= 1e-12 * abs(J0)
    # g^2
    eps = 1e-5 * model.g2
    Jp = forward_J(cfg, make_model(model.m2, model.g2 + eps), fk, pk, PHI0, steps)
    Jm = forward_J(cfg, make_model(model.m2, model.g2 - eps), fk, pk, PHI0, steps)
    fd = (Jp - Jm) / (2 * eps)
    print("g2", dg2, fd)
    assert abs(dg2 - fd) <= 1e-6 * abs(fd) + 1e-16
    # phi0
    eps = 1e-5 * PHI0
    fd = (forward_J(cfg, model, fk, pk, PHI0 + eps, steps) -
          forward_J(cfg, model, fk, pk, PHI0 - eps, steps)) / (2 * eps)
    print("phi0", dphi0, fd)
    assert abs(dphi0 - fd) <= 1e-6 * abs(fd) + 1e-16
    # chi(x) at a few sites
    for site in [(0, 0, 0), (3, 5, 1), (7, 2, 6)]:
        eps = 1e-6
        fd = (forward_J(cfg, model, fk, pk, PHI0, steps, site, eps) -
              forward_J(cfg, model, fk, pk, PHI0, steps, site, -eps)) / (2 * eps)
        print("chi", site, dchi[site], fd)
        assert abs(dchi[site] - fd) <= 1e-6 * abs(fd) + 1e-13


# ---- GW energy fraction ------------------------------------------------------

def forward_gw(cfg, model, fk, pk, u, pu, phi0, steps, projection, dchi_site=None, eps_chi=0.0):
    """rho_GW / rho after `steps` steps of the half-spectrum oracle with GW."""
    N = cfg.N
    fk = fk.copy()
    fk[0, 0, 0, 0] = phi0 * N ** 3
    if dchi_site is not None:
        x = np.zeros((N, N, N))
        x[dchi_site] = eps_chi
        fk[1] += O.fft_r2c_3d(x)
    V0 = 0.5 * model.m2 * phi0 ** 2
    H = math.sqrt((DPHI0 ** 2 / 2 + V0) / 3.0)
    stat = cfg.expansion == PSC_EXPANSION_STATIC
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, 1.0, 0.0 if stat else H, 0.0)
    add0 = o.kernel_a()
    o.load(fk, pk, 0.0, 1.0, 0.0 if stat else H, add0)
    o.load_gw(u, pu)
    o.step(steps)
    _, tot = o.gw_spectrum(projection)
    return float(tot[0])


def setup_gw(integrator, expansion, N=8, g2=2.0e3):
    cfg, model, ic = parity_setup(N=N, nflds=2, integrator=integrator, expansion=expansion,
                                  kIR=4.0, dt=0.01, g2=g2, omega_star=2e-2, f0=(PHI0, 0.0),
                                  fd0=(DPHI0, 0.0), yoshida_exact=1, with_gw=1)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    u, pu = O.generate_gw_ic(cfg, model, ic)
    return cfg, model, fk, pk, u, pu


@pytest.mark.parametrize("integrator,expansion,projection", [(PSC_VV2, PSC_EXPANSION_SELF, 1),
                                                             (PSC_Y4, PSC_EXPANSION_SELF, 0),
                                                             (PSC_VV2, PSC_EXPANSION_STATIC, 1)])
def test_gw_fraction_adjoint_matches_central_differences(integrator, expansion, projection):
    steps = 5
    cfg, model, fk, pk, u, pu = setup_gw(integrator, expansion)
    J, dg2, dphi0, dchi = O.adjoint_gw(cfg, model, fk, pk, u, pu, 1.0, PHI0, DPHI0, 0.0, steps, projection)
    J0 = forward_gw(cfg, model, fk, pk, u, pu, PHI0, steps, projection)
    print("J", J, J0)
    assert abs(J - J0) <= 1e-12 * abs(J0)
    eps = 1e-5 * model.g2
    fd = (forward_gw(cfg, make_model(model.m2, model.g2 + eps), fk, pk, u, pu, PHI0, steps, projection) -
          forward_gw(cfg, make_model(model.m2, model.g2 - eps), fk, pk, u, pu, PHI0, steps, projection)) / (2 * eps)
    print("g2", dg2, fd)
    assert abs(dg2 - fd) <= 1e-6 * abs(fd) + 1e-18
    eps = 1e-5 * PHI0
    fd = (forward_gw(cfg, model, fk, pk, u, pu, PHI0 + eps, steps, projection) -
          forward_gw(cfg, model, fk, pk, u, pu, PHI0 - eps, steps, projection)) / (2 * eps)
    print("phi0", dphi0, fd)
    assert abs(dphi0 - fd) <= 1e-6 * abs(fd) + 1e-18
    for site in [(0, 0, 0), (3, 5, 1), (7, 2, 6)]:
        eps = 1e-6
        fd = (forward_gw(cfg, model, fk, pk, u, pu, PHI0, steps, projection, site, eps) -
              forward_gw(cfg, model, fk, pk, u, pu, PHI0, steps, projection, site, -eps)) / (2 * eps)
        print("chi", site, dchi[site], fd)
        assert abs(dchi[site] - fd) <= 1e-6 * abs(fd) + 1e-15


def forward_J_model(cfg, model, fk, pk, phi0, steps, model_name, dchi_site=None, eps_chi=0.0):
    """forward_J for any built model header: H(phi0) from its potential."""
    N = cfg.N
    fk = fk.copy()
    fk[0, 0, 0, 0] = phi0 * N ** 3
    if dchi_site is not None:
        x = np.zeros((N, N, N))
        x[dchi_site] = eps_chi
        fk[1] += O.fft_r2c_3d(x)
    V0 = O.model_eval(model, cfg.nflds, (phi0, 0.0), model_name=model_name)["terms"].sum()
    H = math.sqrt((DPHI0 ** 2 / 2 + V0) / 3.0)
    o = O.Oracle(cfg, model, model_name=model_name)
    o.load(fk, pk, 0.0, 1.0, H, 0.0)
    add0 = o.kernel_a()
    o.load(fk, pk, 0.0, 1.0, H, add0)
    o.step(steps)
    f, _ = o.fields()
    return float((f[1] ** 2).mean())


@pytest.mark.parametrize("coupling,field", [(PSC_COUPLING_LAMBDA, "lambda_"), (PSC_COUPLING_G2, "g2")])
def test_adjoint_lambda_phi4_matches_central_differences(coupling, field):
    """The oracle adjoint built with the lambda phi^4 model header: dJ/d lambda
    and dJ/dg2 (the model's coupling derivatives), dJ/dphi0 and dJ/dchi(x)
    against central differences."""
    name, steps = "lambda_phi4", 6
    cfg, _, ic = parity_setup(N=8, nflds=2, integrator=PSC_Y4, kIR=4.0, dt=0.01, omega_star=2e-2,
                              f0=(PHI0, 0.0), fd0=(DPHI0, 0.0), yoshida_exact=1)
    model = make_model(m2=0.0, g2=2.0e3, lam=20.0)
    fk, pk, sc = O.generate_ic(cfg, model, ic, model_name=name)
    J, dc, dphi0, dchi = O.adjoint_chi2(cfg, model, fk, pk, 1.0, PHI0, DPHI0, 0.0, steps, coupling=coupling,
                                        model_name=name)
    J0 = forward_J_model(cfg, model, fk, pk, PHI0, steps, name)
    assert abs(J - J0) <= 1e-12 * abs(J0)
    base = getattr(model, field)
    eps = 1e-5 * base

    def moved(v):
        m = make_model(model.m2, model.g2, model.lambda_)
        setattr(m, field, v)
        return forward_J_model(cfg, m, fk, pk, PHI0, steps, name)

    fd = (moved(base + eps) - moved(base - eps)) / (2 * eps)
    print(field, dc, fd)
    assert abs(dc - fd) <= 1e-6 * abs(fd) + 1e-16
    eps = 1e-5 * PHI0
    fd = (forward_J_model(cfg, model, fk, pk, PHI0 + eps, steps, name) -
          forward_J_model(cfg, model, fk, pk, PHI0 - eps, steps, name)) / (2 * eps)
    assert abs(dphi0 - fd) <= 1e-6 * abs(fd) + 1e-16
    site, eps = (3, 5, 1), 1e-6
    fd = (forward_J_model(cfg, model, fk, pk, PHI0, steps, name, site, eps) -
          forward_J_model(cfg, model, fk, pk, PHI0, steps, name, site, -eps)) / (2 * eps)
    assert abs(dchi[site] - fd) <= 1e-6 * abs(fd) + 1e-13
