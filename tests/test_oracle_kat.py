"""Pin the oracle's building blocks against known answers that do not come
from the oracle itself: the Park-Miller KAT, the reference README's example
output, brute-force DFTs and numpy's pocketfft, and closed forms."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import make_config, make_model, make_ic, PSC_EXPANSION_SELF
from tests._util import parity_setup

PM_M = 2147483647


def test_park_miller_minimal_standard_kat():
    # Park & Miller (1988): from state 1, the 10000th state is 1043618065.
    assert O.park_miller_iterate(1, 10000) == 1043618065


def test_park_miller_schrage_matches_direct_modmul():
    # Schrage's factorisation (initialize.cpp:25-26) equals 16807 x mod (2^31-1).
    x = 123456789
    for _ in range(1000):
        y = O.park_miller_iterate(x, 1)
        assert y == (16807 * x) % PM_M
        x = y


def test_rand_uniform_warmup_offset():
    # The first call advances 99 times before returning (initialize.cpp:22-24):
    # the first value is the 100th iterate of the seed.
    d = O.rng_draws(7, 3)
    for k in range(3):
        assert d[k] == O.park_miller_iterate(7, 100 + k) / PM_M


def test_rand_uniform_debug_seed():
    # seed < 1 -> 0.33 for every draw (initialize.cpp:15).
    assert np.all(O.rng_draws(0, 5) == 0.33)


def test_readme_hubble_initialisation():
    """README.md:114-126: phi^2 inflation, phi0 = 11 MPl, dphi0 = -0.815284e-6
    (omegaStar = 1e-6): kinetic 0.332344, potential 60.5, a = 1, ad = 4.50305,
    add = 19.7235 (initialize.cpp:649-670)."""
    cfg = make_config(16, nflds=1, kIR=25.0, dt=1e-3, omegaStar=1e-6, fStar=1.0, MPl=1.0)
    model = make_model(1.0, 0.0)
    fd0 = -0.815284e-6 / 1e-6  # initderivs / fStar / omegaStar
    ic = make_ic(f0=(11.0, 0.0), fd0=(fd0, 0.0), seed=1)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    assert abs(fd0 ** 2 / 2 - 0.332344) < 5e-7
    assert abs(0.5 * 11.0 ** 2 - 60.5) == 0
    assert sc[0] == 1.0
    assert abs(sc[1] - 4.50305) < 5e-6
    assert abs(sc[2] - 19.7235) < 5e-5


def brute_dft3(x):
    N = x.shape[0]
    n = np.arange(N)
    W = np.exp(-2j * np.pi * np.outer(n, n) / N)
    return np.einsum("ai,bj,ck,ijk->abc", W, W, W, x)


@pytest.mark.parametrize("N", [4, 8, 6, 12])
def test_fft_r2c_brute_force(N):
    x = np.random.default_rng(N).standard_normal((N, N, N))
    ref = brute_dft3(x)[:, :, : N // 2 + 1]
    assert np.abs(O.fft_r2c_3d(x) - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("N", [16, 32, 24, 48])
def test_fft_r2c_c2r_vs_pocketfft(N):
    x = np.random.default_rng(N).standard_normal((N, N, N))
    k = O.fft_r2c_3d(x)
    ref = np.fft.rfftn(x)
    assert np.abs(k - ref).max() < 1e-14 * np.abs(ref).max()
    back = O.fft_c2r_3d(ref)
    assert np.abs(back - np.fft.irfftn(ref) * N ** 3).max() < 1e-12 * N ** 3


def test_fft_closed_forms():
    N = 16
    x = np.full((N, N, N), 2.5)
    k = O.fft_r2c_3d(x)
    assert abs(k[0, 0, 0] - 2.5 * N ** 3) < 1e-9
    assert np.abs(k).sum() - abs(k[0, 0, 0]) < 1e-9
    # cos(2 pi (3x + 5y + 2z) / N) -> N^3/2 at (3, 5, 2) (and its conjugate).
    g = np.indices((N, N, N))
    x = np.cos(2 * np.pi * (3 * g[0] + 5 * g[1] + 2 * g[2]) / N)
    k = O.fft_r2c_3d(x)
    assert abs(k[3, 5, 2] - N ** 3 / 2) < 1e-9
    k[3, 5, 2] = 0
    assert np.abs(k).max() < 1e-9


@pytest.mark.parametrize("n", [64, 3, 6, 24, 96, 384])
def test_fft_1d_vs_numpy(n):
    """Power-of-two lengths and 3 * 2^k (one radix-3 step over radix-2 halves)."""
    x = np.random.default_rng(3).standard_normal(n) + 1j * np.random.default_rng(4).standard_normal(n)
    assert np.abs(O.fft_1d(x, -1) - np.fft.fft(x)).max() < 1e-12
    assert np.abs(O.fft_1d(x, +1) - np.fft.ifft(x) * n).max() < 1e-12


@pytest.mark.parametrize("nflds", [1, 2])
def test_ic_hermitian_and_dc(nflds):
    """Initial half spectra describe real fields: f(-k) = conj f(k) in the
    kz = 0, N/2 planes, real corners, and the DC closed form
    (initialize.cpp:752-756)."""
    N = 16
    cfg, model, ic = parity_setup(N=N, nflds=nflds)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    for arr in (fk, pk):
        for kz in (0, N // 2):
            pl = arr[:, :, :, kz]
            mi = (-np.arange(N)) % N
            np.testing.assert_array_equal(pl, np.conj(pl[:, mi][:, :, mi]))
    N3 = N ** 3
    for q in range(nflds):
        assert fk[q, 0, 0, 0] == ic.f0[q] * N3
        assert pk[q, 0, 0, 0] == ic.fd0[q] * N3
    # The c2r of a Hermitian half spectrum is the real field whose r2c gives it back.
    x = O.fft_c2r_3d(fk[0]) / N3
    np.testing.assert_allclose(O.fft_r2c_3d(x), fk[0], atol=1e-12 * np.abs(fk[0]).max())


def test_ic_shell_statistics():
    """<|f_k|^2> = norm^2 / (omega a^2) with norm = (omegaStar/fStar)
    (N/dx)^{3/2} / sqrt 2 (Rayleigh amplitudes, initialize.cpp:181-182)."""
    N, kIR = 32, 25.0
    cfg, model, ic = parity_setup(N=N, nflds=2, kIR=kIR, omega_star=1e-3, seed=11)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    L = 2 * math.pi / kIR
    norm2 = (1e-3 * (N / (L / N)) ** 1.5 / math.sqrt(2)) ** 2
    i = np.arange(N)
    n = np.where(i <= N // 2, i, i - N)
    kz = np.arange(N // 2 + 1)
    p2 = kIR ** 2 * (n[:, None, None] ** 2 + n[None, :, None] ** 2 + kz[None, None, :] ** 2)
    meff = [1.0, 2.5e5 * 1.0 ** 2]
    for q in range(2):
        om = np.sqrt(p2 + meff[q])
        ratio = (np.abs(fk[q]) ** 2) * om / norm2
        ratio = ratio[:, :, 1: N // 2]  # interior planes (independent modes)
        m = ratio.mean()
        assert abs(m - 1.0) < 5.0 / math.sqrt(ratio.size), m


def test_pi_k_relation_of_modes():
    """pi_k = a^{alpha-3} (-omega Im(l - r) - H f)  (initialize.cpp:196-197):
    with H = 0 (no homogeneous energy) |pi_k|^2 averages omega^2 |f_k|^2."""
    N, kIR = 16, 25.0
    cfg, model, ic = parity_setup(N=N, nflds=1, kIR=kIR, omega_star=1e-3, f0=(0.0, 0.0),
                                  fd0=(0.0, 0.0), seed=5, g2=0.0)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    assert sc[3] == 0.0
    i = np.arange(N)
    n = np.where(i <= N // 2, i, i - N)
    kz = np.arange(N // 2 + 1)
    om2 = kIR ** 2 * (n[:, None, None] ** 2 + n[None, :, None] ** 2 + kz[None, None, :] ** 2) + 1.0
    r = (np.abs(pk[0]) ** 2 / om2)[:, :, 1:N // 2].mean() / (np.abs(fk[0]) ** 2)[:, :, 1:N // 2].mean()
    assert abs(r - 1.0) < 0.1


def test_fftw_api_backend_agrees_with_radix2():
    """The FFTW3-API backend of the CPU baseline (libfftw3, else MKL's FFTW3
    interface) against the oracle's pinned radix-2 transforms: one Yoshida-4
    step of a 16^3 two-field lattice."""
    from oracle import oracle as O
    from pspectcosmo_amd.abi import PSC_Y4
    from tests._util import parity_setup, rel_err
    import pytest
    cfg, model, ic = parity_setup(N=16, integrator=PSC_Y4)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    res = []
    for backend in (0, 1):
        name = O.set_fft_backend(backend, 2)
        if name is None:
            pytest.skip("no FFTW3 API (libfftw3 or MKL) on this host")
        o = O.Oracle(cfg, model)
        o.load(fk, pk, 0.0, sc[0], sc[1], sc[4])
        o.step(2)
        res.append(o.get_k())
    O.set_fft_backend(0)
    assert rel_err(res[1][0], res[0][0]) < 1e-13 and rel_err(res[1][1], res[0][1]) < 1e-13


def _visit_order(N):
    """The loop nest of initialize() (initialize.cpp:685-775), DC corner
    included: (i, j, k, real)."""
    h = N // 2
    out = []
    for k in (0, h):
        for i in (0, h):
            out += [(i, j, k, False) for j in range(1, h)]
        for j in (0, h):
            out += [(i, j, k, False) for i in range(1, h)]
        for i in range(1, h):
            out += [(i, j, k, False) for j in range(1, h)]
            out += [(i, j, k, False) for j in range(h + 1, N)]
    for k in range(1, h):
        out += [(i, j, k, False) for i in range(N) for j in range(N)]
    out += [(i, j, k, True) for k in (0, h) for i in (0, h) for j in (0, h)]
    return out


@pytest.mark.parametrize("kcutoff", [0.0, 5.0, 3.2])
def test_initial_modes_decode_to_the_park_miller_stream(kcutoff):
    """Each generated mode is f = L + R, pi = ap (-om Im(L - R) - H Re f) +
    i ap (om Re(L - R) - H Im f) with L = al e^{2 pi i u3}, R = ar e^{2 pi i u4},
    a_{l,r} = rms sqrt(log(1/u_{1,2})) / (sqrt 2 a) (set_mode,
    initialize.cpp:146-220).  Inverting that for every complex mode of both
    fields recovers the four uniforms it used; they must be the Park-Miller
    stream (pinned above) taken four at a time in the visiting order, field
    after field, skipping the modes above kcutoff, which are zero
    (initialize.cpp:138-145).  A wrong visiting order, a cut mode that still
    consumes draws, or a wrong amplitude or phase convention fails here."""
    N, kIR = 16, 1.0
    cfg, model, ic = parity_setup(N=N, kIR=kIR, omega_star=1e-2, g2=50.0, seed=4242, f0=(1.0, 0.3),
                                  fd0=(-0.7, 0.1))
    ic.kcutoff = kcutoff
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    a, H = sc[0], sc[3]
    ap = a ** (cfg.alpha - 3)
    dx = (2 * math.pi / kIR) / N
    norm = cfg.omegaStar / cfg.fStar * (N / dx) ** 1.5 / math.sqrt(2.0)
    meff = O.model_eval(model, 2, [1.0, 0.3])["meff"]
    order = _visit_order(N)
    sidx = lambda i: i if i <= N // 2 else i - N
    k2cut = kcutoff ** 2 if 0 < kcutoff < N * kIR else 0.0
    kept = [not (k2cut > 0 and kIR ** 2 * (sidx(i) ** 2 + sidx(j) ** 2 + k * k) > k2cut) for i, j, k, _ in order]
    per_field = sum(kept)
    if kcutoff > 0:
        assert per_field < len(order)
    draws = O.rng_draws(ic.seed, 4 * per_field * 2)
    checked = 0
    for q in range(2):
        n = 0
        for (i, j, k, real), keep in zip(order, kept):
            f, p = fk[q, i, j, k], pk[q, i, j, k]
            if not keep:
                assert f == 0 and p == 0
                continue
            d = 4 * (q * per_field + n)
            n += 1
            if real:
                continue   # imaginary parts zeroed: not invertible (the DC is overwritten anyway)
            p2 = kIR ** 2 * (sidx(i) ** 2 + sidx(j) ** 2 + k * k)
            om = math.sqrt(p2 + a * a * meff[q]) if p2 + meff[q] > 0 else math.sqrt(p2)
            rms = norm / math.sqrt(om)
            D = complex((p.imag / ap + H * f.imag) / om, -(p.real / ap + H * f.real) / om)
            Lm, Rm = (f + D) / 2, (f - D) / 2
            u = [math.exp(-(abs(Lm) * a * math.sqrt(2.0) / rms) ** 2),
                 math.exp(-(abs(Rm) * a * math.sqrt(2.0) / rms) ** 2),
                 (np.angle(Lm) / (2 * math.pi)) % 1.0, (np.angle(Rm) / (2 * math.pi)) % 1.0]
            for m in range(4):
                e = abs(u[m] - draws[d + m])
                assert min(e, 1 - e) < 1e-9, (q, i, j, k, m, u[m], draws[d + m])
            checked += 1
    assert checked > 2 * 0.5 * per_field
