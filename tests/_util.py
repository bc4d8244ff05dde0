"""Shared configuration builders for the tests.

The configurations named in the benchmark (kIR = 25, dt = 0.002-0.005) violate
the leapfrog stability bound at the highest lattice momentum (k_max dt = 6.9 -
11.1 against 2 for VV and 1.57 for Yoshida-4), so accuracy and parity are
checked on a declared stable variant: the reference's own time-step rule
dt = 2 / (k_max / a) / 10 (pspectcosmo.cpp:213-233, k_max = kIR sqrt(3) N / 2),
and a vacuum amplitude omegaStar / fStar chosen so the chi coupling term is
nonlinear but bounded."""
import math

import numpy as np

from pspectcosmo_amd.abi import (make_config, make_model, make_ic, PSC_VV2, PSC_FP64,
                                 PSC_EXPANSION_SELF)

G2_PREHEAT = 2.5e5


def reference_dt(N, kIR, a=1.0):
    """calculate_dt (pspectcosmo.cpp:213-233) at k = kIR sqrt(3) N / 2."""
    return 2.0 / (kIR * math.sqrt(3.0) * N / 2.0 / a) / 10.0


def parity_setup(N=32, nflds=2, integrator=PSC_VV2, precision=PSC_FP64, omega_star=1e-4,
                 kIR=25.0, dt=None, g2=G2_PREHEAT, m2=1.0, seed=1, batch=1, use_graph=0,
                 expansion=PSC_EXPANSION_SELF, yoshida_exact=0, f0=(1.0, 0.0), fd0=(-0.7, 0.0),
                 with_gw=0, pad=0, n_pad=0):
    if dt is None:
        dt = reference_dt(N, kIR)
    cfg = make_config(N, nflds=nflds, integrator=integrator, precision=precision, batch=batch,
                      kIR=kIR, dt=dt, omegaStar=omega_star, fStar=1.0, MPl=1.0,
                      use_graph=use_graph, expansion=expansion, yoshida_exact=yoshida_exact,
                      with_gw=with_gw, pad=pad, n_pad=n_pad)
    model = make_model(m2, g2 if nflds == 2 else 0.0)
    ic = make_ic(f0=f0, fd0=fd0, seed=seed)
    return cfg, model, ic


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def fluct_rel_err(a, b):
    """Error relative to the fluctuations (DC mode removed) of half spectra."""
    a = np.array(a)
    b = np.array(b)
    a[..., 0, 0, 0] = 0
    b[..., 0, 0, 0] = 0
    return rel_err(a, b)
