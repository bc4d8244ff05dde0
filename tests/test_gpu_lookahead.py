"""psc_step keeps the look-ahead kick across calls and takes its dt on the
device.  Checks: n calls of psc_step(1) leave the state bit-identical to one
psc_step(n) (unpadded, GW, dealiased, ensemble; graphs and eager), also with
reads in between (which restore the synchronised momenta from the sync
buffer); the reference's per-step dt rule with output-time clamps
(calculate_dt, pspectcosmo.cpp:213-233, 622-627) evolved on the device
against the oracle stepping the same rule on the host; a dt change between
calls against the oracle."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import (PSC_VV2, PSC_Y4, PSC_FP64, PSC_FP32, PSC_DT_REFERENCE, PSC_EXPANSION_SELF)
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err

pytestmark = pytest.mark.gpu


def _lattice(cfg, model, fk, pk, sc, gw=None):
    lat = Lattice(cfg)
    for b in range(cfg.batch):
        lat.set_model(b, model)
        lat.load_ic_k(b, fk, pk, sc[0], sc[1])
    if gw is not None:
        lat.load_gw_k(0, *gw)
    return lat


def _state(lat, gw=False):
    out = [np.concatenate([x.ravel() for x in lat.get_k(b)]) for b in range(lat.B)]
    out += [lat.scalars(b) for b in range(lat.B)]
    if gw:
        out += [x.ravel() for x in lat.get_gw_k(0)]
    return out


@pytest.mark.parametrize("precision", [PSC_FP64, PSC_FP32])
@pytest.mark.parametrize("integrator", [PSC_VV2, PSC_Y4])
@pytest.mark.parametrize("use_graph", [0, 1])
def test_single_steps_equal_one_call(precision, integrator, use_graph):
    cfg, model, ic = parity_setup(N=32, integrator=integrator, precision=precision, use_graph=use_graph)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    a = _lattice(cfg, model, fk, pk, sc)
    a.step(6)
    b = _lattice(cfg, model, fk, pk, sc)
    for _ in range(6):
        b.step(1)
    for x, y in zip(_state(a), _state(b)):
        np.testing.assert_array_equal(x, y)
    # reads between the calls restore the synchronised state and do not change it
    c = _lattice(cfg, model, fk, pk, sc)
    e = []
    for _ in range(3):
        c.step(2)
        e.append(c.energies(0))
        c.scalars(0)
    np.testing.assert_array_equal(e[-1], a.energies(0))
    for x, y in zip(_state(a), _state(c)):
        np.testing.assert_array_equal(x, y)


def test_single_steps_gw_padded_and_ensemble():
    cfg, model, ic = parity_setup(N=16, integrator=PSC_Y4, with_gw=1, use_graph=1)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    gw = O.generate_gw_ic(cfg, model, ic)
    a, b = _lattice(cfg, model, fk, pk, sc, gw), _lattice(cfg, model, fk, pk, sc, gw)
    a.step(4)
    for _ in range(4):
        b.step(1)
    for x, y in zip(_state(a, True), _state(b, True)):
        np.testing.assert_array_equal(x, y)
    cfg, model, ic = parity_setup(N=32, integrator=PSC_VV2, pad=2, use_graph=1)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    a, b = _lattice(cfg, model, fk, pk, sc), _lattice(cfg, model, fk, pk, sc)
    a.step(5)
    for _ in range(5):
        b.step(1)
    for x, y in zip(_state(a), _state(b)):
        np.testing.assert_array_equal(x, y)
    cfg, model, ic = parity_setup(N=16, integrator=PSC_VV2, batch=3, use_graph=1, precision=PSC_FP32)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    a, b = _lattice(cfg, model, fk, pk, sc), _lattice(cfg, model, fk, pk, sc)
    a.step(5)
    for _ in range(5):
        b.step(1)
    for x, y in zip(_state(a), _state(b)):
        np.testing.assert_array_equal(x, y)


def _reference_dt(N, kIR, a, t, t_out, cap):
    """calculate_dt and the main loop's clamps (pspectcosmo.cpp:213-233, 622-627)."""
    dt = min(2.0 / (kIR * math.sqrt(3.0) * N / 2.0 / a) / 10.0, cap)
    if t + dt > t_out:
        dt = t_out - t
    return dt


@pytest.mark.parametrize("integrator", [PSC_VV2, PSC_Y4])
def test_adaptive_dt_matches_oracle(integrator):
    """Device-side dt rule against the oracle stepping with the same rule:
    two output intervals (tOutputFreq = 3e-3) of 16^3, FP64, landing on the
    output times; steps past the output time are empty."""
    N, kIR, cap = 16, 25.0, 3e-3
    cfg, model, ic = parity_setup(N=N, integrator=integrator, kIR=kIR, use_graph=1)
    cfg.dt_rule = PSC_DT_REFERENCE
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = _lattice(cfg, model, fk, pk, sc)
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    steps = 0
    for t_out in (cap, 2 * cap):
        lat.set_output_times(t_out, cap)
        lat.step(40)   # more than needed: the rest are empty steps
        while True:
            s = orc.scalars()
            if abs(s[0] - t_out) < 1e-10:
                break
            orc.set_dt(_reference_dt(N, kIR, s[1], s[0], t_out, cap))
            orc.step(1)
            steps += 1
        g = lat.scalars(0)
        print("t", g[0], "oracle", orc.scalars()[0], "steps", lat.step_count(0), steps, flush=True)
        assert abs(g[0] - t_out) < 1e-10
        assert lat.step_count(0) == steps
        gk, gp = lat.get_k(0)
        ok, op = orc.get_k()
        assert rel_err(gk, ok) <= 1e-10 and rel_err(gp, op) <= 1e-10
        np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10)


def test_adaptive_single_steps_equal_one_call():
    """With the rule, one call of many steps equals single-step calls, as
    for fixed dt."""
    cfg, model, ic = parity_setup(N=16, integrator=PSC_Y4, use_graph=1)
    cfg.dt_rule = PSC_DT_REFERENCE
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    a, b = _lattice(cfg, model, fk, pk, sc), _lattice(cfg, model, fk, pk, sc)
    for lat in (a, b):
        lat.set_output_times(1e300, 1e300)
    a.step(7)
    for _ in range(7):
        b.step(1)
    for x, y in zip(_state(a), _state(b)):
        np.testing.assert_array_equal(x, y)


def test_dt_change_between_calls_matches_oracle():
    cfg, model, ic = parity_setup(N=16, integrator=PSC_Y4, use_graph=1)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    lat = _lattice(cfg, model, fk, pk, sc)
    orc = O.Oracle(cfg, model)
    orc.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    for dt, n in ((cfg.dt, 3), (0.7 * cfg.dt, 4), (1.3 * cfg.dt, 2)):
        lat.set_dt(dt)
        orc.set_dt(dt)
        lat.step(n)
        orc.step(n)
    gk, gp = lat.get_k(0)
    ok, op = orc.get_k()
    assert rel_err(gk, ok) <= 1e-10 and rel_err(gp, op) <= 1e-10
    np.testing.assert_allclose(lat.scalars(0), orc.scalars(), rtol=1e-10)
