"""The oracle's snapshot energy densities (kinetic_energy_lattice,
gradient_energy_lattice: evolution.cpp:7-138; potential and total per site:
output.cpp:1614-1625), pinned by closed forms and by independent code:

* plane waves with a Nyquist component along one axis, at a != 1 so that the
  reference's a-powers (a^6/2 on pi^2, a^2/2 on |grad f|^2) are checked;
* the gradient density against the oracle's separate 3D k-space derivative;
* lattice means against the k-space energies (K, V terms) and against G
  summed in k space without the Nyquist components."""
import numpy as np

from oracle import oracle as O
from tests._util import parity_setup


def test_plane_waves_closed_form():
    N, kIR, a, m2, g2 = 16, 2.0, 1.3, 1.7, 40.0
    cfg, model, ic = parity_setup(N=N, kIR=kIR, m2=m2, g2=g2)
    o = O.Oracle(cfg, model)
    x = 2 * np.pi / N * np.indices((N, N, N))
    th = 3 * x[0] - 2 * x[1] + 5 * x[2]
    ph = x[0] + 4 * x[2]
    phi = 0.3 + np.cos(th) + 0.5 * np.cos(N / 2 * x[1])
    chi = -0.2 + 0.7 * np.sin(ph)
    pphi = 0.1 * np.cos(2 * x[2])
    pchi = 0.4 - 0.3 * np.sin(x[0] + x[1])
    fk = np.stack([O.fft_r2c_3d(phi), O.fft_r2c_3d(chi)])
    pk = np.stack([O.fft_r2c_3d(pphi), O.fft_r2c_3d(pchi)])
    o.load(fk, pk, 0.0, a, 0.0, 0.0)
    pot, kin, grad, tot = o.energy_densities()
    # the Nyquist component along y has no derivative
    dphi = [-3 * kIR * np.sin(th), 2 * kIR * np.sin(th), -5 * kIR * np.sin(th)]
    dchi = [0.7 * kIR * np.cos(ph), 0 * ph, 4 * 0.7 * kIR * np.cos(ph)]
    g_exact = a ** 2 / 2 * sum(dphi[d] ** 2 + dchi[d] ** 2 for d in range(3))
    k_exact = a ** 6 / 2 * (pphi ** 2 + pchi ** 2)
    v_exact = 0.5 * m2 * phi ** 2 + 0.5 * g2 * phi ** 2 * chi ** 2
    np.testing.assert_allclose(grad, g_exact, rtol=0, atol=1e-11 * g_exact.max())
    np.testing.assert_allclose(kin, k_exact, rtol=0, atol=1e-13 * k_exact.max())
    np.testing.assert_allclose(pot, v_exact, rtol=0, atol=1e-12 * v_exact.max())
    np.testing.assert_allclose(tot, pot + kin + grad, rtol=1e-15, atol=0)


def test_against_3d_derivative_and_energies():
    N = 16
    cfg, model, ic = parity_setup(N=N, omega_star=2e-2, kIR=4.0, g2=2e3)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
    o.step(4)
    a = o.scalars()[1]
    pot, kin, grad, tot = o.energy_densities()
    g3 = a ** 2 / 2 * sum(o.derivative(q, d) ** 2 for q in range(2) for d in range(3))
    np.testing.assert_allclose(grad, g3, rtol=0, atol=1e-12 * g3.max())
    e = o.energies()
    # mean kinetic density = a^6/2 <pi^2> = a^12 (K_phi + K_chi)
    assert abs(kin.mean() - a ** 12 * (e[0] + e[2])) < 1e-12 * kin.mean()
    assert abs(pot.mean() - (e[4] + e[5])) < 1e-12 * pot.mean()
    f, p = o.fields()
    np.testing.assert_allclose(kin, a ** 6 / 2 * (p ** 2).sum(axis=0), rtol=1e-14)
    # Parseval: the mean gradient density is a^4 times G summed over the
    # half spectrum with the Nyquist component of each n_d dropped
    k, _ = o.get_k()
    i = np.arange(N)
    nn = np.where(i <= N // 2, i, i - N).astype(float)
    nx = np.where(np.abs(nn) == N // 2, 0.0, nn)
    kz = np.arange(N // 2 + 1).astype(float)
    nzz = np.where(kz == N // 2, 0.0, kz)
    n2d = nx[:, None, None] ** 2 + nx[None, :, None] ** 2 + nzz[None, None, :] ** 2
    w = np.full(N // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    Gk = sum(cfg.kIR ** 2 * (w * n2d * np.abs(k[q]) ** 2).sum() / N ** 6 / (2 * a ** 2) for q in range(2))
    assert abs(grad.mean() - a ** 4 * Gk) < 1e-12 * grad.mean()
