"""Loading a member from real-space fields (psc_load_fields, the
continue_run branch of initialize.cpp:590-616): the state read back from one
context and loaded into another continues identically."""
import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import PSC_VV2, PSC_FP32, PSC_FP64
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [(PSC_FP64, 1e-13), (PSC_FP32, 1e-6)])
def test_real_space_image_round_trip(precision, tol):
    cfg, model, ic = parity_setup(N=32, integrator=PSC_VV2, precision=precision)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    a = Lattice(cfg)
    a.set_model(0, model)
    a.load_ic_k(0, fk, pk, sc[0], sc[1])
    a.step(4)
    f, p = a.fields()
    s = a.scalars()
    b = Lattice(cfg)
    b.set_model(0, model)
    b.load_fields(0, f, p, s[0], s[1], s[2])
    ka, kb = a.get_k(), b.get_k()
    print("round trip", rel_err(kb[0], ka[0]), rel_err(kb[1], ka[1]))
    assert rel_err(kb[0], ka[0]) <= tol and rel_err(kb[1], ka[1]) <= tol
    np.testing.assert_allclose(b.scalars(), s, rtol=max(tol, 1e-12))
    a.step(3)
    b.step(3)
    ka, kb = a.get_k(), b.get_k()
    assert rel_err(kb[0], ka[0]) <= 10 * tol and rel_err(kb[1], ka[1]) <= 10 * tol
