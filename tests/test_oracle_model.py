"""Pin the model headers (the model.h interface: potential_energy, dvdf,
d2Vdf2, effective_mass, plus dvdf_delta and the coupling derivatives) that
the HIP library and the oracle are both built with: every derivative against
central differences of the function it differentiates, the fluctuation form
of V'(mean + d) - V'(mean) against the FP64 difference (and its FP32 form
against the FP32 difference it replaces), and the single-field preheat model
against the reference's shipped model.h (model.h:35-98, m = 1)."""
import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import (make_model, PSC_COUPLING_G2, PSC_COUPLING_M2, PSC_COUPLING_LAMBDA)

MODELS = {
    "preheat": (make_model(m2=1.3, g2=2.5e3), [PSC_COUPLING_G2, PSC_COUPLING_M2]),
    "lambda_phi4": (make_model(m2=0.0, g2=2.5e3, lam=7.0), [PSC_COUPLING_G2, PSC_COUPLING_LAMBDA]),
}
POINTS = [(1.1, 0.2), (-0.7, 0.05), (0.3, -0.4)]


def V(name, model, f):
    return O.model_eval(model, 2, f, model_name=name)["terms"].sum()


def grad(fun, x, h):
    x = np.asarray(x, dtype=float)
    g = np.zeros_like(x)
    for i in range(len(x)):
        e = np.zeros_like(x)
        e[i] = h
        # 4th-order central difference
        g[i] = (-fun(x + 2 * e) + 8 * fun(x + e) - 8 * fun(x - e) + fun(x - 2 * e)) / (12 * h)
    return g


@pytest.mark.parametrize("name", list(MODELS))
@pytest.mark.parametrize("f", POINTS)
def test_derivatives_against_differences(name, f):
    model, _ = MODELS[name]
    assert O.model_name(name) == name
    e = O.model_eval(model, 2, f, model_name=name)
    assert e["nterms"] == 2
    dv = grad(lambda x: V(name, model, x), f, 1e-4)
    np.testing.assert_allclose(e["dvdf"], dv, rtol=1e-8, atol=1e-8 * np.abs(dv).max())
    H = np.array([grad(lambda x: O.model_eval(model, 2, x, model_name=name)["dvdf"][q], f, 1e-4)
                  for q in range(2)])
    np.testing.assert_allclose(H, H.T, rtol=1e-7, atol=1e-6)
    hess = e["d2Vdf2"]   # index fld + fld2
    np.testing.assert_allclose([hess[0], hess[1], hess[2]], [H[0, 0], H[0, 1], H[1, 1]], rtol=1e-7,
                               atol=1e-6 * np.abs(H).max())
    # effective masses: the diagonal of the Hessian (model.h:86-98)
    np.testing.assert_allclose(e["meff"], [hess[0], hess[2]], rtol=1e-15)


@pytest.mark.parametrize("name", list(MODELS))
@pytest.mark.parametrize("f", POINTS)
def test_coupling_derivatives(name, f):
    model, couplings = MODELS[name]
    fields = {0: "g2", 1: "m2", 2: "lambda_"}
    for k in couplings:
        e = O.model_eval(model, 2, f, coupling=k, model_name=name)
        assert e["ok"]
        base = getattr(model, fields[k])
        h = 1e-3 * max(abs(base), 1.0)

        def at(c, what):
            m = make_model(model.m2, model.g2, model.lambda_)
            setattr(m, fields[k], c)
            r = O.model_eval(m, 2, f, model_name=name)
            return r["terms"].sum() if what == "V" else r["dvdf"]

        dV = (at(base + h, "V") - at(base - h, "V")) / (2 * h)
        ddv = (at(base + h, "dv") - at(base - h, "dv")) / (2 * h)
        assert abs(e["dV_dc"] - dV) <= 1e-8 * max(abs(dV), 1e-12)
        np.testing.assert_allclose(e["d2V_dfdc"], ddv, rtol=1e-8, atol=1e-10)
    assert not O.model_eval(model, 2, f, coupling=5, model_name=name)["ok"]


@pytest.mark.parametrize("name", list(MODELS))
@pytest.mark.parametrize("f", POINTS)
def test_dvdf_delta(name, f):
    """dvdf_delta(mean, d) = V'(mean + d) - V'(mean); in FP32 the fluctuation
    form keeps its relative accuracy for d ~ 1e-6 |mean|, where the FP32
    difference of the two V' values would have none."""
    model, _ = MODELS[name]
    d = np.array([3e-2, -2e-2])
    e = O.model_eval(model, 2, f, d, model_name=name)
    ref = (O.model_eval(model, 2, np.add(f, d), model_name=name)["dvdf"] -
           O.model_eval(model, 2, f, model_name=name)["dvdf"])
    np.testing.assert_allclose(e["delta"], ref, rtol=1e-12)
    small = np.array([2e-6, -1e-6])
    e = O.model_eval(model, 2, f, small, model_name=name)
    f32 = np.asarray(f, dtype=np.float32).astype(float)
    s32 = small.astype(np.float32).astype(float)
    exact = (O.model_eval(model, 2, f32 + s32, model_name=name)["dvdf"] -
             O.model_eval(model, 2, f32, model_name=name)["dvdf"])
    np.testing.assert_allclose(e["delta32"], exact, rtol=2e-5)


def test_preheat_single_field_is_the_shipped_model():
    """model.h as shipped (nflds = 1, m = 1): V = phi^2 / 2, dvdf = phi,
    d2Vdf2 = 1, effective_mass = 1 (model.h:35-98)."""
    m = make_model(m2=1.0, g2=2.5e5)
    for phi in (11.0, -0.3):
        e = O.model_eval(m, 1, (phi, 5.0))   # chi is ignored with one field
        assert e["nterms"] == 1
        assert e["terms"][0] == 0.5 * phi * phi and e["terms"][1:].sum() == 0.0
        assert e["dvdf"][0] == phi
        assert e["d2Vdf2"][0] == 1.0 and e["meff"][0] == 1.0
