"""The reference's two run-stopping checks, forced from the host so that
every kernel still runs to completion:

* a field whose lattice mean is not a number (meansvars, output.cpp:450-454)
  returns PSC_ERR_NONFINITE from psc_means and psc_energies;
* a negative scale factor (scale, output.cpp:480-483) returns
  PSC_ERR_NEGATIVE_A from psc_get_scalars and psc_energies.

The driver's exit status for both (1, as the reference's exit(1)) is in
test_gpu_driver.py."""
import numpy as np
import pytest

from pspectcosmo_amd.abi import PSC_ERR_NONFINITE, PSC_ERR_NEGATIVE_A, PSC_FP32, PSC_FP64
from pspectcosmo_amd.lattice import Lattice, PscError, generate_ic
from tests._util import parity_setup

pytestmark = pytest.mark.gpu


def fresh(precision):
    cfg, model, ic = parity_setup(N=16, precision=precision, kIR=4.0, g2=2e3, omega_star=2e-2, dt=0.01)
    fk, pk, sc = generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    return lat


@pytest.mark.parametrize("precision", [PSC_FP64, PSC_FP32])
def test_nonfinite_field_is_reported(precision):
    print("case NaN in field 1, precision", precision, flush=True)
    lat = fresh(precision)
    lat.step(2)
    lat.means()   # finite: no error
    f, p = lat.fields()
    s = lat.scalars()
    f[1, 3, 4, 5] = np.nan
    lat.load_fields(0, f, p, s[0], s[1], s[2])
    for call in (lat.means, lat.energies):
        with pytest.raises(PscError) as e:
            call()
        assert e.value.code == PSC_ERR_NONFINITE
        # phi and chi share one packed transform, so a NaN in chi reaches
        # phi's spectrum too and the first field reported is field 0
        assert "not numerical" in str(e.value)
    lat.close()


def test_negative_scale_factor_is_reported():
    print("case a < 0 after one step from da/dt = -1000", flush=True)
    lat = fresh(PSC_FP64)
    s = lat.scalars()
    lat.set_scalars(0, s[0], s[1], -1000.0, s[3])   # a = -9 after the step, the fields stay finite
    lat.step(1)
    for call in (lat.scalars, lat.energies):
        with pytest.raises(PscError) as e:
            call()
        assert e.value.code == PSC_ERR_NEGATIVE_A
    lat.close()
