"""The oracle's dealiased stepper (evolve_VV_k_pad with padding,
extract_padding, kernel_field_k_pad and kernel_a_k_pad, evolution.cpp:1271-1759;
N_pad = 2N and 3N/2), pinned by:

* dealiasing: for fields without Nyquist content, V' = (m2 + g2 chi^2) phi,
  g2 phi^2 chi from the padded grid equals the exact cubic products computed on
  a 4N grid (numpy), mode by mode, folded modulo N_pad (the discrete aliasing
  of a product sampled on the padded grid: nothing folds back into the kept
  modes at 2N, the images at +-N_pad do at 3N/2); modes with an index N/2 read
  the +N/2 copy doubled, as extract_padding does;
* linear V' (g2 = 0): the padded stepper reproduces the unpadded one, in a
  static background with Nyquist content and in an expanding one without it
  (where the padded and unpadded lattice means of V agree);
* time reversal of the nonlinear padded step in a static background;
* kernel_a_k_pad: V averaged over the padded lattice (the exact mean plus, at
  3N/2, the content of V at the multiples of N_pad)."""
import numpy as np
import pytest

from oracle import oracle as O
from pspectcosmo_amd.abi import PSC_VV2, PSC_Y4, PSC_EXPANSION_STATIC, PSC_EXPANSION_SELF
from tests._util import parity_setup, rel_err


def setup(N=16, pad=2, n_pad=0, **kw):
    cfg, model, ic = parity_setup(N=N, pad=pad, n_pad=n_pad, **kw)
    fk, pk, sc = O.generate_ic(cfg, model, ic)
    return cfg, model, fk, pk, sc


def loaded(cfg, model, fk, pk, a, ad):
    o = O.Oracle(cfg, model)
    o.load(fk, pk, 0.0, a, ad, 0.0)
    add = o.kernel_a()
    o.load(fk, pk, 0.0, a, ad, add)
    return o


def band_limited(N, seed):
    """Two random real fields with |n_d| < N/2 in every direction."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(2):
        x = rng.standard_normal((N, N, N))
        k = np.fft.rfftn(x)
        k[N // 2, :, :] = 0
        k[:, N // 2, :] = 0
        k[:, :, N // 2] = 0
        out.append(np.fft.irfftn(k, s=(N, N, N), axes=(0, 1, 2)) * 0.3 + (1.0 if not out else 0.2))
    return out


@pytest.mark.parametrize("n_pad", [0, 24])
def test_dealiased_vprime_equals_exact_products(n_pad):
    N, m2, g2 = 16, 1.3, 7.0
    Np = n_pad or 2 * N
    cfg, model, ic = parity_setup(N=N, pad=2, n_pad=n_pad, m2=m2, g2=g2, expansion=PSC_EXPANSION_STATIC)
    phi, chi = band_limited(N, 3)
    fk = np.stack([np.fft.rfftn(phi), np.fft.rfftn(chi)])
    o = O.Oracle(cfg, model)
    o.load(fk, np.zeros_like(fk), 0.0, 1.0, 0.0, 0.0)
    got = o.vprime_k()
    # exact products on a 4N grid (cubic terms of modes below N/2 stay below 3N/2)
    M = 4 * N
    def up(k):
        big = np.zeros((M, M, M // 2 + 1), dtype=complex)
        h = N // 2
        for sx in (slice(0, h), slice(M - h, M)):
            for sy in (slice(0, h), slice(M - h, M)):
                src_x = slice(0, h) if sx.start == 0 else slice(N - h, N)
                src_y = slice(0, h) if sy.start == 0 else slice(N - h, N)
                big[sx, sy, :h] = k[src_x, src_y, :h]
        return np.fft.irfftn(big, s=(M, M, M), axes=(0, 1, 2)) * (M / N) ** 3
    P, X = up(fk[0]), up(fk[1])
    exact = [np.fft.fftn((m2 + g2 * X ** 2) * P) / (M / N) ** 3, np.fft.fftn(g2 * P ** 2 * X) / (M / N) ** 3]
    # the padded grid samples the product: its spectrum is the exact one summed
    # over the images s + m N_pad (only m = 0 is nonzero at N_pad = 2N)
    imgs = [-Np, 0, Np]
    folded_any = 0.0
    for q in range(2):
        for i in range(N):
            si = i if i <= N // 2 else i - N
            ci = 2.0 if i == N // 2 else 1.0
            for j in range(N):
                sj = j if j <= N // 2 else j - N
                cj = 2.0 if j == N // 2 else 1.0
                for k in range(N // 2 + 1):
                    ck = 2.0 if k == N // 2 else 1.0
                    img = sum(exact[q][(si + a) % M, (sj + b) % M, (k + c) % M]
                              for a in imgs for b in imgs for c in imgs if (a, b, c) != (0, 0, 0))
                    folded_any = max(folded_any, abs(img))
                    ref = ci * cj * ck * (exact[q][si % M, sj % M, k] + img)
                    assert abs(got[q, i, j, k] - ref) <= 1e-9 * np.abs(exact[q]).max(), (q, i, j, k)
    # images fold into the kept modes at 3N/2 only
    if Np == 2 * N:
        assert folded_any <= 1e-12 * np.abs(exact[0]).max()
    else:
        assert folded_any > 1e-4 * np.abs(exact[0]).max()
    # the unpadded force aliases the same products
    cfg0, _, _ = parity_setup(N=N, pad=0, m2=m2, g2=g2, expansion=PSC_EXPANSION_STATIC)
    o0 = O.Oracle(cfg0, model)
    o0.load(fk, np.zeros_like(fk), 0.0, 1.0, 0.0, 0.0)
    assert rel_err(o0.vprime_k(), got) > 1e-3


@pytest.mark.parametrize("integrator,n_pad", [(PSC_VV2, 0), (PSC_Y4, 0), (PSC_Y4, 24)])
def test_linear_force_padded_equals_unpadded_static(integrator, n_pad):
    kw = dict(N=16, g2=0.0, m2=2.0, omega_star=2e-2, kIR=4.0, integrator=integrator,
              expansion=PSC_EXPANSION_STATIC)
    cfg, model, fk, pk, sc = setup(pad=2, n_pad=n_pad, **kw)
    cfg0, _, _, _, _ = setup(pad=0, **kw)
    a = loaded(cfg, model, fk, pk, 1.0, 0.0)
    b = loaded(cfg0, model, fk, pk, 1.0, 0.0)
    a.step(5)
    b.step(5)
    (af, ap), (bf, bp) = a.get_k(), b.get_k()
    assert rel_err(af, bf) < 1e-13 and rel_err(ap, bp) < 1e-13


@pytest.mark.parametrize("n_pad", [0, 24])
def test_linear_force_padded_equals_unpadded_expanding_without_nyquist(n_pad):
    kw = dict(N=16, g2=0.0, m2=2.0, omega_star=2e-2, kIR=4.0, integrator=PSC_Y4, expansion=PSC_EXPANSION_SELF)
    cfg, model, fk, pk, sc = setup(pad=2, n_pad=n_pad, **kw)
    cfg0, _, _, _, _ = setup(pad=0, **kw)
    for arr in (fk, pk):
        arr[:, 8, :, :] = 0
        arr[:, :, 8, :] = 0
        arr[:, :, :, 8] = 0
    a = loaded(cfg, model, fk, pk, sc[0], sc[1])
    b = loaded(cfg0, model, fk, pk, sc[0], sc[1])
    assert abs(a.scalars()[3] - b.scalars()[3]) < 1e-12 * abs(b.scalars()[3])
    a.step(5)
    b.step(5)
    (af, ap), (bf, bp) = a.get_k(), b.get_k()
    assert rel_err(af, bf) < 1e-12 and rel_err(ap, bp) < 1e-12
    np.testing.assert_allclose(a.scalars(), b.scalars(), rtol=1e-12)


@pytest.mark.parametrize("n_pad", [0, 24])
def test_nonlinear_padded_step_is_time_reversible(n_pad):
    kw = dict(N=16, g2=2e3, omega_star=2e-2, kIR=4.0, integrator=PSC_Y4, expansion=PSC_EXPANSION_STATIC)
    cfg, model, fk, pk, sc = setup(pad=2, n_pad=n_pad, **kw)
    o = loaded(cfg, model, fk, pk, 1.0, 0.0)
    f0, p0 = o.get_k()
    o.step(3)
    f1, _ = o.get_k()
    assert rel_err(f1, f0) > 1e-6
    o.set_dt(-cfg.dt)
    o.step(3)
    f2, p2 = o.get_k()
    assert rel_err(f2, f0) < 1e-12 and rel_err(p2, p0) < 1e-12


@pytest.mark.parametrize("n_pad", [0, 24])
def test_kernel_a_averages_v_over_the_padded_lattice(n_pad):
    N, m2, g2 = 16, 1.0, 50.0
    Np = n_pad or 2 * N
    cfg, model, ic = parity_setup(N=N, pad=2, n_pad=n_pad, m2=m2, g2=g2)
    phi, chi = band_limited(N, 5)
    fk = np.stack([np.fft.rfftn(phi), np.fft.rfftn(chi)])
    o = O.Oracle(cfg, model)
    a = 1.0
    o.load(fk, np.zeros_like(fk), 0.0, a, 0.0, 0.0)
    # band-limited fields: V = m2 phi^2 / 2 + g2 phi^2 chi^2 / 2 has content below
    # 2N, which a 4N grid holds exactly; the mean over the padded grid is the sum
    # of V's coefficients at the multiples of N_pad (the exact spatial mean
    # alone at 2N, where no multiple holds content)
    M = 4 * N
    def up(x):
        k = np.fft.fftn(x)
        big = np.zeros((M, M, M), dtype=complex)
        h = N // 2
        idx = np.r_[0:h, M - h:M]
        src = np.r_[0:h, N - h:N]
        big[np.ix_(idx, idx, idx)] = k[np.ix_(src, src, src)]
        return np.fft.ifftn(big).real * (M / N) ** 3
    P, X = up(phi), up(chi)
    Vk = np.fft.fftn(0.5 * m2 * P ** 2 + 0.5 * g2 * P ** 2 * X ** 2) / M ** 3
    imgs = [-Np, 0, Np]
    V = sum(Vk[a % M, b % M, c % M] for a in imgs for b in imgs for c in imgs).real
    exact_mean = Vk[0, 0, 0].real
    if Np == 2 * N:
        assert abs(V - exact_mean) < 1e-13 * abs(exact_mean)
    else:
        assert abs(V - exact_mean) > 1e-9 * abs(exact_mean)
    add = o.kernel_a()
    # static pi = 0: add = (1/3) (alpha + 1) V a^{1 + 2 alpha} with alpha = 0, plus -2K = 0 and G not used
    assert abs(add - V / 3.0) < 1e-12 * V
