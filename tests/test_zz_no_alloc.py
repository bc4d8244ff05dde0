"""psc_energy_densities_noalloc, the entry point that allocates no device
memory, against psc_energy_densities on the same lattice: bit for bit, on two
calls in a row, the second on a smaller lattice (the staging array keeps the
first call's contents, which nothing may read)."""
import numpy as np
import pytest

from pspectcosmo_amd.abi import PSC_FP32, PSC_FP64, PSC_Y4
from pspectcosmo_amd.lattice import Lattice, generate_ic
from tests._util import parity_setup

pytestmark = pytest.mark.gpu


def lattice(N, precision):
    # the reference's dt rule at kIR = 25 keeps the 256^3 step stable
    cfg, model, ic = parity_setup(N=N, precision=precision, integrator=PSC_Y4)
    fk, pk, sc = generate_ic(cfg, model, ic)
    lat = Lattice(cfg)
    lat.set_model(0, model)
    lat.load_ic_k(0, fk, pk, sc[0], sc[1])
    lat.step(3)
    return lat


def test_noalloc_densities_match_the_allocating_call():
    # 256^3: four blocks of 64 y rows through the 32 MiB array; then 64^3 FP32 in one block
    for N, precision in ((256, PSC_FP64), (64, PSC_FP32)):
        print(f"case energy densities noalloc vs allocating, {N}^3 precision {precision}", flush=True)
        lat = lattice(N, precision)
        ref = lat.energy_densities()
        got = lat.energy_densities_noalloc()
        for name, a, b in zip(("pot", "kin", "grad", "tot"), got, ref):
            assert np.all(np.isfinite(b)) and np.abs(b).max() > 0, name
            np.testing.assert_array_equal(a, b, err_msg=name)
        again = lat.energy_densities_noalloc()
        for a, b in zip(again, ref):
            np.testing.assert_array_equal(a, b)
        lat.close()
