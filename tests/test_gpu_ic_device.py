"""Initial conditions generated on the device (psc_generate_ic_device,
psc_generate_gw_ic_device) against the host generator, which matches the
oracle bit for bit: the same modes from the same Park-Miller draws
(initialize.cpp:128-220, 341-470, 685-775), so the spectra agree to a few
units in the last place (the device's log, sin and cos are not glibc's); the
background scalars and d2a/dt2 agree likewise.  Cases: FP64 and FP32, one and
two fields, static and expanding, GW, an ensemble with one seed per member,
the seed < 1 debug stream, and a short evolution from either start."""
import numpy as np
import pytest

from pspectcosmo_amd.abi import (PSC_FP32, PSC_FP64, PSC_Y4, PSC_EXPANSION_STATIC, make_ic)
from pspectcosmo_amd.lattice import Lattice, generate_ic, generate_gw_ic
from tests._util import parity_setup, rel_err

pytestmark = pytest.mark.gpu


def both(cfg, model, ic, gw=False):
    """(host-loaded lattice, device-generated lattice, host scalars, device scalars)."""
    fk, pk, sc = generate_ic(cfg, model, ic)
    h = Lattice(cfg)
    h.set_model(0, model)
    h.load_ic_k(0, fk, pk, sc[0], sc[1])
    d = Lattice(cfg)
    d.set_model(0, model)
    sd = d.generate_ic_device(0, ic)
    if gw:
        u, pu = generate_gw_ic(cfg, model, ic)
        h.load_gw_k(0, u, pu)
        d.generate_gw_ic_device(0, ic)
    return h, d, sc, sd


def max_rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision,nflds,expansion", [
    (PSC_FP64, 2, None), (PSC_FP64, 1, None), (PSC_FP64, 2, PSC_EXPANSION_STATIC), (PSC_FP32, 2, None)])
def test_device_ic_matches_host(precision, nflds, expansion):
    kw = {} if expansion is None else {"expansion": expansion}
    cfg, model, ic = parity_setup(N=32, nflds=nflds, precision=precision, omega_star=2e-2, kIR=4.0, g2=2e3,
                                  seed=11, **kw)
    print("case device IC 32^3 precision", precision, "nflds", nflds, "expansion", cfg.expansion, flush=True)
    h, d, sc, sd = both(cfg, model, ic)
    (hf, hp), (df, dp) = h.get_k(), d.get_k()
    tol = 1e-14 if precision == PSC_FP64 else 1e-7
    ef, ep = max_rel(df, hf), max_rel(dp, hp)
    print(f"field {ef:.2e} momentum {ep:.2e} (max error / max value)")
    assert ef <= tol and ep <= tol
    assert rel_err(df, hf) <= tol and rel_err(dp, hp) <= tol
    np.testing.assert_array_equal(sd, sc[:4])
    np.testing.assert_allclose(d.scalars(), h.scalars(), rtol=1e-12 if precision == PSC_FP64 else 1e-6, atol=0)
    h.step(3)
    d.step(3)
    assert rel_err(d.get_k()[0], h.get_k()[0]) <= 10 * tol
    h.close()
    d.close()


def test_device_gw_ic_matches_host():
    cfg, model, ic = parity_setup(N=32, precision=PSC_FP64, omega_star=2e-2, kIR=4.0, g2=2e3, seed=5, with_gw=1,
                                  integrator=PSC_Y4)
    print("case device GW IC 32^3 fp64", flush=True)
    h, d, _, _ = both(cfg, model, ic, gw=True)
    (hu, hpu), (du, dpu) = h.get_gw_k(), d.get_gw_k()
    eu, ep = max_rel(du, hu), max_rel(dpu, hpu)
    print(f"u {eu:.2e} pi_u {ep:.2e}")
    assert eu <= 1e-14 and ep <= 1e-14
    assert np.all(du[:, 0, 0, 0] == 0) and np.all(dpu[:, 0, 0, 0] == 0)
    h.step(2)
    d.step(2)
    assert rel_err(d.get_gw_k()[0], h.get_gw_k()[0]) <= 1e-13
    h.close()
    d.close()


def test_ensemble_members_with_their_own_seeds():
    B = 3
    cfg, model, _ = parity_setup(N=32, precision=PSC_FP64, omega_star=2e-2, kIR=4.0, g2=2e3, batch=B)
    print("case device IC ensemble of 3, 32^3", flush=True)
    lat = Lattice(cfg)
    for b in range(B):
        lat.set_model(b, model)
        lat.generate_ic_device(b, make_ic(f0=(1.0, 0.0), fd0=(-0.7, 0.0), seed=100 + b))
    cfg1, _, _ = parity_setup(N=32, precision=PSC_FP64, omega_star=2e-2, kIR=4.0, g2=2e3)
    for b in range(B):
        fk, pk, _ = generate_ic(cfg1, model, make_ic(f0=(1.0, 0.0), fd0=(-0.7, 0.0), seed=100 + b))
        gk, gp = lat.get_k(b)
        assert max_rel(gk, fk) <= 1e-14 and max_rel(gp, pk) <= 1e-14
    lat.close()


def test_debug_seed_constant_stream():
    cfg, model, ic = parity_setup(N=16, precision=PSC_FP64, omega_star=2e-2, kIR=4.0, g2=2e3, seed=0)
    h, d, _, _ = both(cfg, model, ic)
    assert max_rel(d.get_k()[0], h.get_k()[0]) <= 1e-14
    h.close()
    d.close()


def test_rejections():
    cfg, model, ic = parity_setup(N=16, precision=PSC_FP64)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    with pytest.raises(Exception):
        lat.generate_gw_ic_device(0, ic)       # no GW evolution in this context
    with pytest.raises(Exception):
        lat.generate_ic_device(1, ic)          # no such member
    lat.close()


def test_first_generation_after_another_context_at_512():
    """A fresh context's zero fills and copies are ordered with its kernels:
    the first device ICs generated right after creation (on memory another
    context used and freed) equal a second generation.  With the fills on the
    null stream, the fill of pi overlapped the IC kernel at 512^3 and zeroed
    part of the modes it had written."""
    cfg, model, ic = parity_setup(N=512, with_gw=1, integrator=PSC_Y4, precision=PSC_FP32)
    print("case device IC 512^3 GW after another context", flush=True)
    first = Lattice(cfg)
    first.set_model(0, model)
    first.generate_ic_device(0, make_ic(f0=(1.3, 0.1), fd0=(0.2, 0.0), seed=7))
    first.step(1)
    first.close()
    lat = Lattice(cfg)
    lat.set_model(0, model)
    sums = []
    for _ in range(2):
        lat.generate_ic_device(0, ic)
        lat.generate_gw_ic_device(0, ic)
        s, _ = lat.gw_spectrum(0, 0)
        sums.append((lat.spectra(0, 0).sum(), lat.spectra(1, 0).sum(), s.sum(), lat.scalars(0)[3]))
    print("first", sums[0], "second", sums[1], flush=True)
    assert sums[0] == sums[1]
    lat.close()


@pytest.mark.parametrize("nflds,N,seed", [(2, 32, 11), (1, 16, 3), (2, 64, 7)])
def test_device_ic_matches_oracle(nflds, N, seed):
    """The device generator against the oracle's restatement of initialize()
    directly (not through the host generator): modes, background scalars,
    and three steps of both from their own starts."""
    from oracle import oracle as O
    cfg, model, ic = parity_setup(N=N, nflds=nflds, precision=PSC_FP64, omega_star=2e-2, kIR=4.0, g2=2e3,
                                  seed=seed)
    print(f"case device IC vs oracle {N}^3 nflds {nflds} seed {seed}", flush=True)
    ofk, opk, osc = O.generate_ic(cfg, model, ic)
    d = Lattice(cfg)
    d.set_model(0, model)
    sd = d.generate_ic_device(0, ic)
    dfk, dpk = d.get_k()
    ef, ep = max_rel(dfk, ofk), max_rel(dpk, opk)
    print(f"device vs oracle: field {ef:.2e} momentum {ep:.2e} (max error / max value)")
    assert ef <= 1e-14 and ep <= 1e-14
    np.testing.assert_allclose(sd, osc[:4], rtol=1e-15, atol=0)
    s = d.scalars()
    np.testing.assert_allclose(s[1:3], osc[0:2], rtol=1e-15, atol=0)
    np.testing.assert_allclose(s[3], osc[4], rtol=1e-12, atol=0)   # d2a/dt2 from the device's own sums
    o = O.Oracle(cfg, model)
    o.load(ofk, opk, 0.0, osc[0], osc[1], osc[4])
    o.step(3)
    d.step(3)
    gk, gp = d.get_k()
    ok, op = o.get_k()
    print(f"after 3 steps: field {rel_err(gk, ok):.2e} momentum {rel_err(gp, op):.2e}")
    assert rel_err(gk, ok) <= 1e-12 and rel_err(gp, op) <= 1e-12
    np.testing.assert_allclose(d.scalars(), o.scalars(), rtol=1e-12)
    d.close()


def test_device_gw_ic_matches_oracle():
    from oracle import oracle as O
    cfg, model, ic = parity_setup(N=32, precision=PSC_FP64, omega_star=2e-2, kIR=4.0, g2=2e3, seed=5, with_gw=1,
                                  integrator=PSC_Y4)
    print("case device GW IC vs oracle 32^3 fp64", flush=True)
    ou, opu = O.generate_gw_ic(cfg, model, ic)
    d = Lattice(cfg)
    d.set_model(0, model)
    d.generate_ic_device(0, ic)
    d.generate_gw_ic_device(0, ic)
    du, dpu = d.get_gw_k()
    eu, ep = max_rel(du, ou), max_rel(dpu, opu)
    print(f"u {eu:.2e} pi_u {ep:.2e}")
    assert eu <= 1e-14 and ep <= 1e-14
    d.close()


@pytest.mark.parametrize("N,kcutoff,precision", [(32, 300.0, PSC_FP64), (64, 800.0, PSC_FP64), (64, 800.0, PSC_FP32)])
def test_device_ic_with_kcutoff_matches_oracle(N, kcutoff, precision):
    """The kcutoff low-pass on the device: each work item finds its draw
    offset from the kept modes before it (closed-form row counts and the
    interior row table), against the oracle's sequential stream; scalar and
    GW modes."""
    from oracle import oracle as O
    cfg, model, ic = parity_setup(N=N, precision=precision, seed=9, with_gw=1, integrator=PSC_Y4)
    ic.kcutoff = kcutoff
    print(f"case device IC with kcutoff {kcutoff} at {N}^3 precision {precision}", flush=True)
    ofk, opk, osc = O.generate_ic(cfg, model, ic)
    ou, opu = O.generate_gw_ic(cfg, model, ic)
    d = Lattice(cfg)
    d.set_model(0, model)
    d.generate_ic_device(0, ic)
    d.generate_gw_ic_device(0, ic)
    dfk, dpk = d.get_k()
    du, dpu = d.get_gw_k()
    tol = 1e-14 if precision == PSC_FP64 else 1e-7
    errs = [max_rel(dfk, ofk), max_rel(dpk, opk), max_rel(du, ou), max_rel(dpu, opu)]
    print("field, momentum, u, pi_u:", ["%.2e" % e for e in errs])
    assert max(errs) <= tol
    cut = ofk[0] == 0
    assert np.all(np.abs(dfk[0][cut]) <= tol * np.abs(ofk).max())
    d.close()
