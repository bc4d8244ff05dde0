"""Batched ensemble: members with their own seeds, couplings and initial
inflaton values advance in one context; each member and the ensemble-mean
spectra match the FP64 oracle run member by member."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import PSC_VV2, make_model, make_ic
from pspectcosmo_amd.lattice import Lattice
from tests._util import parity_setup, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def members(n):
    with open(os.path.join(ROOT, "pspectcosmo_amd", "workloads", "ensemble_128_64.json")) as fh:
        return json.load(fh)["members"][:n]


def test_ensemble_members_and_mean_spectra_match_oracle():
    B, N, steps = 6, 32, 5
    cfg, _, _ = parity_setup(N=N, batch=B, integrator=PSC_VV2, use_graph=1)
    lat = Lattice(cfg)
    oracles = []
    for m in members(B):
        model = make_model(1.0, m["g2"])
        ic = make_ic(f0=(m["phi0"], 0.0), fd0=(-0.7, 0.0), seed=m["seed"])
        fk, pk, sc = O.generate_ic(cfg, model, ic)
        lat.set_model(m["member"], model)
        lat.load_ic_k(m["member"], fk, pk, sc[0], sc[1])
        c1, _, _ = parity_setup(N=N, batch=1, integrator=PSC_VV2)
        o = O.Oracle(c1, model)
        o.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
        oracles.append(o)
    lat.step(steps)
    mean_gpu = [0.0, 0.0]
    mean_orc = [0.0, 0.0]
    for b, o in enumerate(oracles):
        o.step(steps)
        gk, gp = lat.get_k(b)
        ok, op = o.get_k()
        assert rel_err(gk, ok) <= 1e-10 and rel_err(gp, op) <= 1e-10, b
        np.testing.assert_allclose(lat.scalars(b), o.scalars(), rtol=1e-10)
        for f in range(2):
            mean_gpu[f] = mean_gpu[f] + lat.spectra(f, b) / B
            mean_orc[f] = mean_orc[f] + o.spectra(f) / B
    for f in range(2):
        np.testing.assert_allclose(mean_gpu[f], mean_orc[f], rtol=1e-10, atol=0)


def test_member_groups_match_single_member_runs():
    """At 128^3 FP32 the force chain runs on groups of 8 members (about 128 MiB
    of work array each): with 10 members the groups are 8 and 2.  Every
    member ends bit-identical to its own single-member context, graph replay
    and eager launches alike."""
    B, N, steps = 10, 128, 3
    from pspectcosmo_amd.abi import PSC_FP32
    for use_graph in (1, 0):
        cfg, _, _ = parity_setup(N=N, batch=B, integrator=PSC_VV2, precision=PSC_FP32, use_graph=use_graph)
        print("case member groups 10 x 128^3 FP32, graph", use_graph, flush=True)
        lat = Lattice(cfg)
        ms = members(B)
        for m in ms:
            lat.set_model(m["member"], make_model(1.0, m["g2"]))
            lat.generate_ic_device(m["member"], make_ic(f0=(m["phi0"], 0.0), fd0=(-0.7, 0.0), seed=m["seed"]))
        lat.step(steps)
        c1, _, _ = parity_setup(N=N, batch=1, integrator=PSC_VV2, precision=PSC_FP32, use_graph=use_graph)
        for b in (0, 7, 8, 9):
            m = ms[b]
            one = Lattice(c1)
            one.set_model(0, make_model(1.0, m["g2"]))
            one.generate_ic_device(0, make_ic(f0=(m["phi0"], 0.0), fd0=(-0.7, 0.0), seed=m["seed"]))
            one.step(steps)
            (gk, gp), (sk, sp) = lat.get_k(b), one.get_k(0)
            assert np.array_equal(gk, sk) and np.array_equal(gp, sp), b
            assert np.array_equal(lat.scalars(b), one.scalars(0)), b
            one.close()
        lat.close()
