// psc_oracle_adjoint.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU reference gradient of end-of-run observables of the spectral
// lattice step: the discrete adjoint of evolve_VV_k_nopad
// (evolution.cpp:1131-1198) including the Friedmann coupling through
// kernel_a_k_nopad (evolution.cpp:1061-1090), with the two real fields packed
// as P = FFT(phi + i chi) over the full N^3 spectrum.  It records every
// substep of the forward run (no checkpointing: test sizes are small) and
// sweeps back.  Pinned by central finite differences of the half-spectrum
// forward restatement in psc_oracle.cpp (tests/test_oracle_adjoint.py).
//
// Parameters: g^2; phi0 = the homogeneous inflaton value with the fluctuation
// modes held fixed (it sets the DC mode phi0 N^3 and the initial expansion rate
// da/dt = a0 H(phi0), initialize.cpp:649-668); the initial real-space chi
// field.  d2a/dt2 at t = 0 is recomputed from the state (initialize.cpp:797).
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#include <omp.h>

#include "../include/psc.h"

typedef std::complex<double> cd;

namespace {

const double kPi = 3.14159265358979323846264338327950288;

struct Fft1 {
    int n;
    std::vector<cd> tw;
    std::vector<int> rev;
    explicit Fft1(int n_) : n(n_), tw(n_ / 2), rev(n_) {
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        for (int k = 0; k < n / 2; ++k) tw[k] = std::polar(1.0, -2.0 * kPi * k / n);
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b)
                if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            rev[i] = r;
        }
    }
    void exec(cd* a, int sign) const {
        for (int i = 0; i < n; ++i)
            if (rev[i] > i) std::swap(a[i], a[rev[i]]);
        for (int len = 2; len <= n; len <<= 1)
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; ++k) {
                    cd w = tw[k * (n / len)];
                    if (sign > 0) w = std::conj(w);
                    cd u = a[i + k], v = a[i + k + len / 2] * w;
                    a[i + k] = u + v;
                    a[i + k + len / 2] = u - v;
                }
    }
};

// Unnormalised 3D c2c transform in place (sign -1 forward F, +1 inverse F^H).
void fft3(std::vector<cd>& x, int N, int sign, const Fft1& f) {
    const long NN = (long)N * N;
#pragma omp parallel
    {
        std::vector<cd> buf(N);
#pragma omp for
        for (long l = 0; l < NN; ++l) {
            for (int z = 0; z < N; ++z) buf[z] = x[l * N + z];
            f.exec(buf.data(), sign);
            for (int z = 0; z < N; ++z) x[l * N + z] = buf[z];
        }
#pragma omp for
        for (long l = 0; l < NN; ++l) {
            const long xi = l / N, z = l % N;
            for (int y = 0; y < N; ++y) buf[y] = x[(xi * N + y) * N + z];
            f.exec(buf.data(), sign);
            for (int y = 0; y < N; ++y) x[(xi * N + y) * N + z] = buf[y];
        }
#pragma omp for
        for (long l = 0; l < NN; ++l) {
            const long y = l / N, z = l % N;
            for (int xi = 0; xi < N; ++xi) buf[xi] = x[(xi * NN) + y * N + z];
            f.exec(buf.data(), sign);
            for (int xi = 0; xi < N; ++xi) x[(xi * NN) + y * N + z] = buf[xi];
        }
    }
}

inline int sidx(int i, int N) { return i <= N / 2 ? i : i - N; }
inline double pw2(double x) { return x * x; }

struct Adj {
    psc_config cfg;
    double m2, g2;
    int N;
    long n3;
    double N3;
    Fft1 f1;
    std::vector<double> n2;  // |n|^2 per packed index
    std::vector<double> w;
    bool stat;
    Adj(const psc_config& c, double m2_, double g2_) : cfg(c), m2(m2_), g2(g2_), N(c.N), f1(c.N) {
        n3 = (long)N * N * N;
        N3 = (double)n3;
        n2.resize(n3);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k)
                    n2[((long)i * N + j) * N + k] = pw2(sidx(i, N)) + pw2(sidx(j, N)) + pw2(sidx(k, N));
        if (c.integrator == PSC_Y4) {
            if (c.yoshida_exact) {
                const double cr = std::cbrt(2.0);
                w = {1 / (2 - cr), -cr / (2 - cr), 1 / (2 - cr)};
            } else {
                w = {1.351207191959657771818, -1.702414403875838200264, 1.351207191959657771818};
            }
        } else {
            w = {1.0};
        }
        stat = c.expansion == PSC_EXPANSION_STATIC;
    }
    // r = F^H f / N^3 (packed real space)
    std::vector<cd> real_space(const std::vector<cd>& f) const {
        std::vector<cd> r = f;
        fft3(r, N, +1, f1);
        for (auto& v : r) v /= N3;
        return r;
    }
    cd vprime(cd r) const {
        const double p = r.real(), q = r.imag();
        return cd(m2 * p + g2 * p * q * q, g2 * p * p * q);
    }
    // F(f, a) = -a^{1+al} kIR^2 n^2 f - a^{3+al} F(V'(r))
    std::vector<cd> force(const std::vector<cd>& f, double a) const {
        std::vector<cd> r = real_space(f);
        for (auto& v : r) v = vprime(v);
        fft3(r, N, -1, f1);
        const double c1 = std::pow(a, 1 + cfg.alpha) * pw2(cfg.kIR), c3 = std::pow(a, 3 + cfg.alpha);
        std::vector<cd> F(n3);
        for (long i = 0; i < n3; ++i) F[i] = -c1 * n2[i] * f[i] - c3 * r[i];
        return F;
    }
    struct KGV {
        double K, G, V, Vg;  // Vg = <phi^2 chi^2 / 2>
    };
    KGV kgv(const std::vector<cd>& f, const std::vector<cd>& p, double a) const {
        double sk = 0, sg = 0, sv = 0, svg = 0;
        for (long i = 0; i < n3; ++i) {
            sk += std::norm(p[i]);
            sg += n2[i] * std::norm(f[i]);
        }
        std::vector<cd> r = real_space(f);
        for (auto& v : r) {
            const double ph = v.real(), ch = v.imag();
            sv += 0.5 * m2 * ph * ph + 0.5 * g2 * ph * ph * ch * ch;
            svg += 0.5 * ph * ph * ch * ch;
        }
        KGV o;
        o.K = 0.5 * sk / std::pow(a, 6) / (N3 * N3);
        o.G = pw2(cfg.kIR) * sg / N3 / 2.0 / (a * a) / N3;
        o.V = sv / N3;
        o.Vg = svg / N3;
        return o;
    }
    double cA(double a) const { return 1.0 / 3.0 * pw2(cfg.fStar / cfg.MPl) * std::pow(a, 1 + 2 * cfg.alpha); }
    double A(const KGV& e, double a) const {
        if (stat) return 0.0;
        const int al = cfg.alpha;
        return cA(a) * ((al - 2) * e.K + al * e.G + (al + 1) * e.V);
    }

    struct Rec {
        std::vector<cd> f, p, ph;
        double a, b, add, ah, bh;
    };
    // One substep with weight wp; fills rec with its start state and pi_h.
    void substep(double wp, std::vector<cd>& f, std::vector<cd>& p, double& a, double& b, double& add,
                 Rec& rec) const {
        const double h = wp * cfg.dt / 2.0;
        rec.f = f;
        rec.p = p;
        rec.a = a;
        rec.b = b;
        rec.add = add;
        const double bh = b + h * add;
        std::vector<cd> F1 = force(f, a);
        std::vector<cd> ph(n3);
        for (long i = 0; i < n3; ++i) ph[i] = p[i] + h * F1[i];
        const double ah = stat ? 1.0 : a + bh * h;
        const double c = 2 * h * std::pow(ah, cfg.alpha - 3);
        for (long i = 0; i < n3; ++i) f[i] += c * ph[i];
        a = stat ? 1.0 : ah + bh * h;
        std::vector<cd> F2 = force(f, a);
        for (long i = 0; i < n3; ++i) p[i] = ph[i] + h * F2[i];
        add = A(kgv(f, p, a), a);
        b = stat ? 0.0 : bh + h * add;
        rec.ph = std::move(ph);
        rec.ah = ah;
        rec.bh = bh;
    }

    // Adjoint of the force F(f, a) at (f, a) applied to Fbar, together with an
    // extra real-space adjoint rV (from the V part of kernel_a): adds to fbar,
    // returns the contributions to abar and g2bar.
    void force_adjoint(const std::vector<cd>& f, double a, const std::vector<cd>& Fbar,
                       const std::vector<cd>* rV, std::vector<cd>& fbar, double& abar, double& g2bar) const {
        const int al = cfg.alpha;
        const double c1 = std::pow(a, 1 + al) * pw2(cfg.kIR), c3 = std::pow(a, 3 + al);
        double da = 0;
        for (long i = 0; i < n3; ++i) {
            fbar[i] += -c1 * n2[i] * Fbar[i];
            da += (std::conj(Fbar[i]) * (-(1 + al) * std::pow(a, al) * pw2(cfg.kIR) * n2[i] * f[i])).real();
        }
        std::vector<cd> vb(n3);
        for (long i = 0; i < n3; ++i) vb[i] = -c3 * Fbar[i];
        fft3(vb, N, +1, f1);  // vbar = F^H(-a^{3+al} Fbar)
        std::vector<cd> r = real_space(f);
        std::vector<cd> q(n3);
        double dv = 0, dg = 0;
        for (long i = 0; i < n3; ++i) {
            const double ph = r[i].real(), ch = r[i].imag();
            const cd v = vprime(r[i]);
            dv += vb[i].real() * v.real() + vb[i].imag() * v.imag();
            dg += vb[i].real() * ph * ch * ch + vb[i].imag() * ph * ph * ch;
            const double hpp = m2 + g2 * ch * ch, hpc = 2 * g2 * ph * ch, hcc = g2 * ph * ph;
            q[i] = cd(hpp * vb[i].real() + hpc * vb[i].imag(), hpc * vb[i].real() + hcc * vb[i].imag());
            if (rV) q[i] += (*rV)[i];
        }
        fft3(q, N, -1, f1);
        for (long i = 0; i < n3; ++i) fbar[i] += q[i] / N3;
        abar += da + (3 + al) / a * dv;
        g2bar += dg;
    }

    // Adjoint of add = A(f, p, a): adds to fbar, pbar; returns the real-space
    // adjoint of the V part; contributions to abar, g2bar.
    std::vector<cd> A_adjoint(const std::vector<cd>& f, const std::vector<cd>& p, double a, double addbar,
                              std::vector<cd>& fbar, std::vector<cd>& pbar, double& abar, double& g2bar) const {
        std::vector<cd> rV(n3, 0.0);
        if (stat || addbar == 0.0) return rV;
        const int al = cfg.alpha;
        const KGV e = kgv(f, p, a);
        const double c = cA(a);
        const double Q = (al - 2) * e.K + al * e.G + (al + 1) * e.V;
        const double ck = addbar * c * (al - 2) / std::pow(a, 6) / (N3 * N3);
        const double cg = addbar * c * al * pw2(cfg.kIR) / (a * a) / (N3 * N3);
        for (long i = 0; i < n3; ++i) {
            pbar[i] += ck * p[i];
            fbar[i] += cg * n2[i] * f[i];
        }
        std::vector<cd> r = real_space(f);
        const double cv = addbar * c * (al + 1) / N3;
        for (long i = 0; i < n3; ++i) rV[i] = cv * vprime(r[i]);
        abar += addbar * c * ((1 + 2 * al) * Q / a - 6.0 * (al - 2) * e.K / a - 2.0 * al * e.G / a);
        g2bar += addbar * c * (al + 1) * e.Vg;
        return rV;
    }
};

// Packed full spectrum from FFTW half spectra of nflds fields.
void pack(int N, int nflds, const double* hk, std::vector<cd>& P) {
    const int Nh = N / 2 + 1;
    const long nk = (long)N * N * Nh;
    P.assign((long)N * N * N, 0.0);
    const cd* h = reinterpret_cast<const cd*>(hk);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                cd v[2] = {0.0, 0.0};
                for (int q = 0; q < nflds; ++q) {
                    if (k <= N / 2) v[q] = h[q * nk + ((long)i * N + j) * Nh + k];
                    else v[q] = std::conj(h[q * nk + ((long)((N - i) % N) * N + (N - j) % N) * Nh + (N - k)]);
                }
                P[((long)i * N + j) * N + k] = v[0] + cd(0, 1) * v[1];
            }
}

}  // namespace

extern "C" {

// Gradient of J = <chi^2> at the end of nsteps steps (objective 0).  Inputs:
// initial half spectra, a0, the homogeneous phi0, dphi0, chi0 used for
// H(phi0).  Outputs: J, dJ/dg2, dJ/dphi0, dJ/dchi(x) (N^3, may be NULL).
int orc_adjoint_chi2(const psc_config* cfg, const psc_model* model, const double* f_k, const double* pi_k,
                     double a0, double phi0, double dphi0, double chi0, int nsteps, double* J,
                     double* g2bar_out, double* phi0bar_out, double* chibar_out) {
    if (cfg->nflds != 2) return PSC_ERR_ARG;
    Adj s(*cfg, model->m2, model->g2);
    const long n3 = s.n3;
    const double N3 = s.N3;
    std::vector<cd> f, p;
    pack(cfg->N, 2, f_k, f);
    pack(cfg->N, 2, pi_k, p);
    // background at t = 0 (initialize.cpp:649-668) and kernel_a on the state
    const double V0 = 0.5 * s.m2 * phi0 * phi0 + 0.5 * s.g2 * phi0 * phi0 * chi0 * chi0;
    const double f2m = pw2(cfg->fStar / cfg->MPl);
    const double H = std::sqrt((dphi0 * dphi0 / 2.0 + V0) * f2m / 3.0);
    double a = a0, b = s.stat ? 0.0 : H * a0;
    const std::vector<cd> f0 = f, p0 = p;
    double add = s.A(s.kgv(f, p, a), a);
    std::vector<Adj::Rec> recs;
    const int S = (int)s.w.size();
    recs.resize((size_t)nsteps * S);
    for (int st = 0; st < nsteps; ++st)
        for (int q = 0; q < S; ++q) s.substep(s.w[q], f, p, a, b, add, recs[(size_t)st * S + q]);
    // objective and its adjoint at the end
    std::vector<cd> r = s.real_space(f);
    double jj = 0;
    for (long i = 0; i < n3; ++i) jj += pw2(r[i].imag());
    jj /= N3;
    std::vector<cd> fbar(n3), pbar(n3, 0.0), Fpend(n3, 0.0);
    for (long i = 0; i < n3; ++i) fbar[i] = cd(0, 2.0 * r[i].imag() / N3);
    fft3(fbar, cfg->N, -1, s.f1);
    for (auto& v : fbar) v /= N3;
    double abar = 0, bbar = 0, addbar = 0, g2bar = 0;
    // end state of the substep being undone
    std::vector<cd> fe = f, pe = p;
    double ae = a;
    for (long sidx_ = (long)recs.size() - 1; sidx_ >= 0; --sidx_) {
        const Adj::Rec& R = recs[sidx_];
        const double h = s.w[sidx_ % S] * cfg->dt / 2.0;
        // b' = bh + h add'
        addbar += h * bbar;
        double bhbar = bbar;
        // add' = A(f', p', a')
        std::vector<cd> rV = s.A_adjoint(fe, pe, ae, addbar, fbar, pbar, abar, g2bar);
        // p' = ph + h F2;  merged force adjoint at (f', a')
        std::vector<cd> phbar = pbar;
        std::vector<cd> Fb(n3);
        for (long i = 0; i < n3; ++i) Fb[i] = h * pbar[i] + Fpend[i];
        s.force_adjoint(fe, ae, Fb, &rV, fbar, abar, g2bar);
        if (s.stat) abar = 0;
        // a' = ah + bh h
        double ahbar = abar;
        bhbar += h * abar;
        // f' = f + 2h ah^{al-3} ph
        const double c = 2 * h * std::pow(R.ah, cfg->alpha - 3);
        double dot = 0;
        for (long i = 0; i < n3; ++i) {
            phbar[i] += c * fbar[i];
            dot += (std::conj(fbar[i]) * R.ph[i]).real();
        }
        ahbar += 2 * h * (cfg->alpha - 3) * std::pow(R.ah, cfg->alpha - 4) * dot;
        // ah = a + bh h
        abar = s.stat ? 0.0 : ahbar;
        bhbar += h * ahbar;
        // ph = p + h F1 (F1 deferred to the start point)
        pbar = phbar;
        for (long i = 0; i < n3; ++i) Fpend[i] = h * phbar[i];
        // bh = b + h add
        bbar = s.stat ? 0.0 : bhbar;
        addbar = s.stat ? 0.0 : h * bhbar;
        fe = R.f;
        pe = R.p;
        ae = R.a;
    }
    // t = 0: the pending force and add0 = A(f0, p0, a0)
    std::vector<cd> rV0 = s.A_adjoint(f0, p0, a0, addbar, fbar, pbar, abar, g2bar);
    s.force_adjoint(f0, a0, Fpend, &rV0, fbar, abar, g2bar);
    *J = jj;
    *g2bar_out = g2bar;
    // phi0: DC mode phi0 N^3 and da/dt(0) = a0 H(phi0), dH/dphi0 = f2m V'(phi0) / (6 H)
    const double vphi = s.m2 * phi0 + s.g2 * phi0 * chi0 * chi0;
    *phi0bar_out = N3 * fbar[0].real() + (s.stat ? 0.0 : bbar * a0 * f2m * vphi / (6.0 * H));
    if (chibar_out) {
        std::vector<cd> g = fbar;
        fft3(g, cfg->N, +1, s.f1);
        for (long i = 0; i < n3; ++i) chibar_out[i] = g[i].imag();
    }
    (void)pbar;
    return 0;
}

}  // extern "C"
