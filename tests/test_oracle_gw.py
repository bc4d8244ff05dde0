"""Pin the oracle's GW parts: polarisation tensors, projector identities,
the free GW wave equation and the source of a single plane wave."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import PSC_GW_PROJ_REFERENCE, PSC_GW_PROJ_TT, PSC_EXPANSION_STATIC
from tests._util import parity_setup, rel_err

CI = [0, 1, 2, 0, 0, 1]
CJ = [0, 1, 2, 1, 2, 2]


def kvecs(N):
    i = np.arange(N)
    n = np.where(i <= N // 2, i, i - N)
    kz = np.arange(N // 2 + 1)
    return n[:, None, None], n[None, :, None], kz[None, None, :]


def full_tensor(u):
    """(6, N, N, Nh) symmetric components -> (3, 3, N, N, Nh)."""
    T = np.zeros((3, 3) + u.shape[1:], dtype=u.dtype)
    for c in range(6):
        T[CI[c], CJ[c]] = u[c]
        T[CJ[c], CI[c]] = u[c]
    return T


def test_gw_initial_modes_transverse_traceless():
    """e = uu - vv + uv + vu with u, v orthonormal and orthogonal to k
    (initialize.cpp:66-122): k_i u_ij = 0 and u_ii = 0 for every mode."""
    N = 16
    cfg, model, ic = parity_setup(N=N, with_gw=1, omega_star=1e-3)
    u, pu = O.generate_gw_ic(cfg, model, ic)
    T = full_tensor(u)
    nx, ny, nz = kvecs(N)
    nz_full = np.broadcast_to(nz, T.shape[2:]).copy()
    # stored conjugate indices in the kz = 0, N/2 planes carry the conjugate of
    # the visited mode: transversality holds with the visited mode's k, which
    # differs only by sign, so k_i u_ij = 0 for all stored entries.
    kk = [np.broadcast_to(nx, T.shape[2:]), np.broadcast_to(ny, T.shape[2:]), nz_full]
    # The conjugate stored for a mode with a Nyquist index (N/2) sits at a lattice
    # vector that is not -k (N/2 maps to +N/2), so only non-Nyquist rows are checked.
    ok = (np.abs(kk[0]) != N // 2) & (np.abs(kk[1]) != N // 2) & (kk[2] != N // 2)
    for j in range(3):
        div = sum(kk[i] * T[i, j] for i in range(3))
        assert np.abs(div[ok]).max() < 1e-10 * np.abs(T).max()
    tr = T[0, 0] + T[1, 1] + T[2, 2]
    assert np.abs(tr).max() < 1e-12 * np.abs(T).max()
    assert np.all(u[:, 0, 0, 0] == 0)


def single_mode_state(N, kidx, tensor):
    u = np.zeros((6, N, N, N // 2 + 1), dtype=np.complex128)
    for c in range(6):
        u[c][kidx] = tensor[CI[c], CJ[c]]
    return u


@pytest.mark.parametrize("kind,ref_expected,tt_expected", [
    ("tt", 1.0, 1.0),            # pure TT tensor: both projections keep it
    ("longitudinal", 0.0, 0.0),  # k^ k^
    ("trace", 0.0, 0.0),         # delta_ij: one-sided P delta = P, |P|^2 - |tr P|^2/2 = 2 - 2
    ("vector", 0.5, 0.0),        # (k^ e + e k^)/sqrt2: the one-sided P.u keeps e k^/sqrt2, Lambda removes it
])
def test_gw_projection_closed_forms(kind, ref_expected, tt_expected):
    N = 8
    cfg, model, ic = parity_setup(N=N, nflds=2, with_gw=1, omega_star=0.0)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    kidx = (1, 2, 1)
    kv = np.array([1.0, 2.0, 1.0])
    kh = kv / np.linalg.norm(kv)
    e1 = np.cross(kh, [0, 0, 1.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(kh, e1)
    if kind == "tt":
        t = (np.outer(e1, e1) - np.outer(e2, e2)) / math.sqrt(2)
    elif kind == "longitudinal":
        t = np.outer(kh, kh)
    elif kind == "trace":
        t = np.eye(3)
    else:
        t = (np.outer(kh, e1) + np.outer(e1, kh)) / math.sqrt(2)
    pu = single_mode_state(N, kidx, t)
    o.load_gw(np.zeros_like(pu), pu)
    for proj, expected in ((PSC_GW_PROJ_REFERENCE, ref_expected), (PSC_GW_PROJ_TT, tt_expected)):
        out, tot = o.gw_spectrum(proj)
        # all of sum_ij |m|^2 of the mode sits in bin round(|n|) with weight 2
        b = int(math.floor(math.sqrt(6.0) + 0.5))
        row = out[b - 1]
        hd2 = row[2] / (b / N ** 6 / 4.0 / sc[0] ** 6 / o.energies()[6])
        frob = float((t * t).sum())
        assert abs(hd2 - 2 * expected * frob) < 1e-9 * max(1.0, frob), (kind, proj, hd2)


def test_free_gw_exact_discrete_solution():
    """No fields, static background: every u_ij mode obeys the exact discrete
    VV map of u'' = -k^2 u."""
    N, kIR, dt, n = 8, 1.0, 0.05, 20
    cfg, model, ic = parity_setup(N=N, nflds=1, with_gw=1, kIR=kIR, dt=dt, g2=0.0, m2=1.0,
                                  omega_star=0.0, f0=(0.0, 0.0), fd0=(0.0, 0.0),
                                  expansion=PSC_EXPANSION_STATIC)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, 1.0, 0.0, 0.0)
    cfg2, _, ic2 = parity_setup(N=N, with_gw=1, omega_star=1e-2, seed=3)
    u0, pu0 = O.generate_gw_ic(cfg2, model, ic2)
    o.load_gw(u0, pu0)
    o.step(n)
    u, pu = o.get_gw()
    nx, ny, nz = kvecs(N)
    w2 = kIR ** 2 * (nx ** 2 + ny ** 2 + nz ** 2)
    A = 1 - w2 * dt * dt / 2
    C = -w2 * dt * (1 - w2 * dt * dt / 4)
    ue, pe = u0.copy(), pu0.copy()
    for _ in range(n):
        ue, pe = A * ue + dt * pe, C * ue + A * pe
    assert rel_err(u, ue) < 1e-12 and rel_err(pu, pe) < 1e-12


def test_plane_wave_source_has_no_tt_part():
    """phi = A cos(k.x): S_ij = k_i k_j A^2 sin^2(k.x) lives at 0 and 2k with a
    tensor along k^ k^, so after a step from u = 0 the TT spectrum of pi_u vanishes
    while the source did drive the longitudinal part."""
    N, kIR = 16, 1.0
    cfg, model, ic = parity_setup(N=N, nflds=1, with_gw=1, kIR=kIR, dt=0.01, g2=0.0, m2=1.0,
                                  omega_star=0.0, f0=(0.0, 0.0), fd0=(0.0, 0.0),
                                  expansion=PSC_EXPANSION_STATIC)
    g = np.indices((N, N, N))
    x = 2 * np.pi / N * g
    phi = 0.3 * np.cos(2 * x[0] + 1 * x[1] + 3 * x[2])
    fk = O.fft_r2c_3d(phi)[None]
    o = O.Oracle(cfg, model)
    o.load(fk, np.zeros_like(fk), 0.0, 1.0, 0.0, 0.0)
    z = np.zeros((6, N, N, N // 2 + 1), dtype=np.complex128)
    o.load_gw(z, z)
    o.step(1)
    u, pu = o.get_gw()
    assert np.abs(pu).max() > 1e-3
    out_tt, _ = o.gw_spectrum(PSC_GW_PROJ_TT)
    out_ref, _ = o.gw_spectrum(PSC_GW_PROJ_REFERENCE)
    # Scale of the unprojected kinetic spectrum in the same units: the source
    # mode 2k sits in bin round(2|k|) = 7.
    e_tot = o.energies()[6]
    frob = float((np.abs(full_tensor(pu)) ** 2).sum(axis=(0, 1)).max())
    unproj = 7 / N ** 6 / 4.0 / e_tot * 2 * frob
    assert unproj > 0
    assert np.abs(out_tt[:, 2]).max() < 1e-12 * unproj
    assert np.abs(out_ref[:, 2]).max() < 1e-12 * unproj
