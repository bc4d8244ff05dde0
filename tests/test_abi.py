"""C ABI checks that need no GPU: the library loads, exports every function
include/psc.h declares, and its host-side initial conditions equal the
oracle's restatement of initialize() bit for bit."""
import os
import re
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd import lattice as L
from pspectcosmo_amd.abi import make_config, make_model, make_ic, PscConfig, PscModel, PscIcParams
from tests._util import parity_setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "psc.h")).read()
    return sorted(set(re.findall(r"\b(psc_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(L.exported_symbols())


@pytest.mark.parametrize("model_name", ["preheat", "lambda_phi4"])
def test_library_exports_every_header_symbol(model_name):
    lib = C.CDLL(L.lib_path(model_name))
    missing = [s for s in header_functions() if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.psc_abi_version() == 2
    lib.psc_model_name.restype = C.c_char_p
    assert lib.psc_model_name().decode() == model_name


def test_struct_sizes_match_header():
    # psc_config: 16 ints + 8 doubles; psc_model: 8 doubles; psc_ic_params: 5 doubles + 2 ints + 4 doubles
    assert C.sizeof(PscConfig) == 16 * 4 + 8 * 8
    assert C.sizeof(PscModel) == 64
    assert C.sizeof(PscIcParams) == 5 * 8 + 2 * 4 + 4 * 8


def test_abi_version():
    assert L.lib().psc_abi_version() == 2
    assert L.lib().psc_num_bins(64) == 56


@pytest.mark.parametrize("N,nflds,seed", [(16, 2, 1), (32, 1, 7), (16, 2, 0), (64, 2, 3), (48, 2, 4), (96, 1, 2)])
def test_host_ic_matches_oracle_bitwise(N, nflds, seed):
    cfg, model, ic = parity_setup(N=N, nflds=nflds, seed=seed)
    fk, pk, sc = L.generate_ic(cfg, model, ic)
    ofk, opk, osc = O.generate_ic(cfg, model, ic)
    np.testing.assert_array_equal(fk, ofk)
    np.testing.assert_array_equal(pk, opk)
    np.testing.assert_array_equal(sc, osc[:4])


@pytest.mark.parametrize("nflds", [1, 2])
def test_host_ic_matches_oracle_bitwise_lambda_phi4(nflds):
    """Both sides built with the lambda phi^4 model header: the effective
    masses 3 lambda phi0^2 (+ g2 chi0^2), g2 phi0^2 set the mode frequencies."""
    cfg, _, ic = parity_setup(N=16, nflds=nflds, seed=5, f0=(1.2, 0.1))
    model = make_model(m2=0.0, g2=3.0e3, lam=50.0)
    fk, pk, sc = L.generate_ic(cfg, model, ic, model_name="lambda_phi4")
    ofk, opk, osc = O.generate_ic(cfg, model, ic, model_name="lambda_phi4")
    np.testing.assert_array_equal(fk, ofk)
    np.testing.assert_array_equal(pk, opk)
    np.testing.assert_array_equal(sc, osc[:4])


@pytest.mark.parametrize("N", [16, 64])
def test_host_gw_ic_matches_oracle_bitwise(N):
    cfg, model, ic = parity_setup(N=N, nflds=2, with_gw=1)
    u, pu = L.generate_gw_ic(cfg, model, ic)
    ou, opu = O.generate_gw_ic(cfg, model, ic)
    np.testing.assert_array_equal(u, ou)
    np.testing.assert_array_equal(pu, opu)


def test_create_rejects_bad_config():
    cfg = make_config(40)  # neither 2^k nor 3 * 2^k
    h = C.c_void_p()
    assert L.lib().psc_create(C.byref(cfg), C.byref(h)) == 1


@pytest.mark.parametrize("N,kcutoff", [(16, 150.0), (32, 300.0), (64, 800.0)])
def test_host_ic_with_kcutoff_matches_oracle_bitwise(N, kcutoff):
    """The kcutoff low-pass (initialize.cpp:138-145): the host generator's
    row-parallel skip-ahead over the kept modes against the oracle's
    sequential restatement, scalar fields and GW modes."""
    cfg, model, ic = parity_setup(N=N, nflds=2, seed=9, with_gw=1)
    ic.kcutoff = kcutoff
    fk, pk, sc = L.generate_ic(cfg, model, ic)
    ofk, opk, osc = O.generate_ic(cfg, model, ic)
    cut = np.count_nonzero(ofk[0] == 0)
    assert 0.2 * ofk[0].size < cut < 0.9 * ofk[0].size   # kIR = 25: a real part of the modes is cut
    np.testing.assert_array_equal(fk, ofk)
    np.testing.assert_array_equal(pk, opk)
    u, pu = L.generate_gw_ic(cfg, model, ic)
    ou, opu = O.generate_gw_ic(cfg, model, ic)
    np.testing.assert_array_equal(u, ou)
    np.testing.assert_array_equal(pu, opu)
