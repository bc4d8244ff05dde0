// psc_ic_device.hpp -- initial conditions generated on the device, straight
// into a member's packed spectra: the modes of initialize() (set_mode,
// initialize.cpp:128-220, visited as initialize.cpp:685-775) and of
// initialize_GW (initialize.cpp:273-470).  One work item per visited mode;
// each positions its own Park-Miller stream at the mode's first draw by
// modular exponentiation, so every mode gets the draws of the sequential
// order.  The arithmetic is the host generator's, in FP64 with no
// contraction; only the device's log, sin and cos differ from the host's in
// the last bits.
#pragma once

#include "psc_common.hpp"
#include "psc_fft.hpp"
#include "psc_ic_modes.hpp"

namespace psc {

// Per-call constants, computed on the host (psc_generate_ic's background and
// normalisation).
struct IcConsts {
    int N, nflds, seed, pad;
    unsigned long long first[2];   // first draw of each field's stream (GW: one stream)
    double norm, dp2, a, ap, H, N3;
    double meff[2], f0[2], fd0[2];
    double kIR;
    psc_ic::KeptIndex kept;        // the kcutoff low-pass; kept.rowpre is a device table
};

// Draw offset / 4 of mode m (work item t): its visiting position, or with a
// cutoff the number of kept modes visited before it.
__device__ __forceinline__ unsigned long long ic_rank(const IcConsts& p, bool dc, const psc_ic::Mode& m,
                                                      unsigned long long t) {
    return p.kept.cut.on() ? p.kept.rank(p.N, dc, m, t) : m.v;
}

__device__ __forceinline__ int ic_sidx(int i, int N) { return i <= N / 2 ? i : i - N; }

// Amplitudes of one mode from its four draws (set_mode, initialize.cpp:181-199):
// f = rl + rr + i (il + ir), pi = ap (-om (il - ir) - H f.re) + i ap (om (rl - rr) - H f.im).
__device__ inline void ic_amplitudes(psc_ic::Stream st, double rms, double a, double& rl, double& il, double& rr,
                                     double& ir) {
    PSC_NO_CONTRACT
    const double u1 = st.uniform();
    st.next();
    const double u2 = st.uniform();
    st.next();
    const double u3 = st.uniform();
    st.next();
    const double u4 = st.uniform();
    const double al = rms / sqrt(2.0) * sqrt(log(1.0 / u1)) / a;
    const double ar = rms / sqrt(2.0) * sqrt(log(1.0 / u2)) / a;
    double sn, cs;
    sincos(2.0 * 3.14159265358979323846264338327950288 * u3, &sn, &cs);
    rl = al * cs;
    il = al * sn;
    sincos(2.0 * 3.14159265358979323846264338327950288 * u4, &sn, &cs);
    rr = ar * cs;
    ir = ar * sn;
}

// Scalar fields: P_f = phi_k + i chi_k and P_pi likewise at k and at -k
// (conjugates), the DC mode set to the homogeneous values times N^3.
template <typename T>
__global__ void __launch_bounds__(256)
k_ic_fields(cpx<T>* __restrict__ Pf, cpx<T>* __restrict__ Pp, IcConsts p, size_t PS, unsigned long long count) {
    PSC_NO_CONTRACT
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int N = p.N;
    const psc_ic::Mode m = psc_ic::mode_at(N, true, t);
    const int px = ic_sidx(m.i, N), py = ic_sidx(m.j, N);
    const size_t at = (size_t)m.i * PS + (size_t)m.j * N + m.k;
    const size_t cj = (size_t)((N - m.i) % N) * PS + (size_t)((N - m.j) % N) * N + (N - m.k) % N;
    if (!p.kept.cut.keeps((int64_t)px * px + (int64_t)py * py + (int64_t)m.k * m.k)) {   // cut: zero, no draws
        Pf[at] = Pp[at] = cpx<T>{(T)0, (T)0};
        if (!m.real) Pf[cj] = Pp[cj] = cpx<T>{(T)0, (T)0};
        return;
    }
    const unsigned long long rank = ic_rank(p, true, m, t);
    const double p2 = p.dp2 * ((double)px * (double)px + (double)py * (double)py + (double)m.k * (double)m.k);
    double fr[2] = {0.0, 0.0}, fi[2] = {0.0, 0.0}, pr[2] = {0.0, 0.0}, pim[2] = {0.0, 0.0};
    for (int q = 0; q < p.nflds; ++q) {
        const double om = (p2 + p.meff[q] > 0.0) ? sqrt(p2 + (p.a * p.a) * p.meff[q]) : sqrt(p2);
        const double rms = om > 0.0 ? p.norm / sqrt(om) : 0.0;
        double rl, il, rr, ir;
        ic_amplitudes(psc_ic::Stream::at(p.seed, p.first[q] + 4 * rank), rms, p.a, rl, il, rr, ir);
        fr[q] = rl + rr;
        fi[q] = il + ir;
        pr[q] = p.ap * (-om * (il - ir) - p.H * fr[q]);
        pim[q] = p.ap * (om * (rl - rr) - p.H * fi[q]);
        if (m.real) fi[q] = pim[q] = 0.0;
    }
    if (m.real && m.i == 0 && m.j == 0 && m.k == 0) {   // initialize.cpp:776-790
        for (int q = 0; q < p.nflds; ++q) {
            fr[q] = p.f0[q] * p.N3;
            pr[q] = p.ap * p.fd0[q] * p.N3;
            fi[q] = pim[q] = 0.0;
        }
    }
    Pf[at] = cpx<T>{(T)(fr[0] - fi[1]), (T)(fi[0] + fr[1])};
    Pp[at] = cpx<T>{(T)(pr[0] - pim[1]), (T)(pim[0] + pr[1])};
    if (!m.real) {
        Pf[cj] = cpx<T>{(T)(fr[0] + fi[1]), (T)(fr[1] - fi[0])};
        Pp[cj] = cpx<T>{(T)(pr[0] + pim[1]), (T)(pr[1] - pim[0])};
    }
}

// GW modes (initialize_GW, initialize.cpp:355-470): the six components
// u_ij = e_ij f, pi_u = ap e_ij (-kk (il - ir) - H u.re, kk (rl - rr) - H u.im)
// packed in pairs (11 + i 22, 33 + i 12, 13 + i 23) into the three arrays
// U + a * astride.  No DC mode.
template <typename T>
__global__ void __launch_bounds__(256)
k_ic_gw(cpx<T>* __restrict__ U, cpx<T>* __restrict__ PU, size_t astride, IcConsts p, size_t PS,
        unsigned long long count) {
    PSC_NO_CONTRACT
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int N = p.N;
    const psc_ic::Mode m = psc_ic::mode_at(N, false, t);
    const int px = ic_sidx(m.i, N), py = ic_sidx(m.j, N);
    // a cut mode stays zero (the arrays are cleared before the launch) and takes no draws
    if (!p.kept.cut.keeps((int64_t)px * px + (int64_t)py * py + (int64_t)m.k * m.k)) return;
    const double kv[3] = {px * p.kIR, py * p.kIR, m.k * p.kIR};
    double e[3][3];
    psc_ic::polarisation(kv, e);
    const double kk = sqrt((p.kIR * p.kIR) * ((double)px * (double)px + (double)py * (double)py + (double)m.k * (double)m.k));
    const double rms = p.norm / sqrt(kk);
    double rl, il, rr, ir;
    ic_amplitudes(psc_ic::Stream::at(p.seed, p.first[0] + 4 * ic_rank(p, false, m, t)), rms, p.a, rl, il, rr, ir);
    const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
    double ur[6], ui[6], pr[6], pim[6];
    for (int c = 0; c < 6; ++c) {
        const double ec = e[ci[c]][cj[c]];
        ur[c] = ec * (rl + rr);
        const double ui0 = ec * (il + ir);
        pr[c] = p.ap * (ec * (-kk * (il - ir) - p.H * ur[c]));
        const double pi0 = p.ap * (ec * (kk * (rl - rr) - p.H * ui0));
        ui[c] = m.real ? 0.0 : ui0;
        pim[c] = m.real ? 0.0 : pi0;
    }
    const size_t at = (size_t)m.i * PS + (size_t)m.j * N + m.k;
    const size_t cjx = (size_t)((N - m.i) % N) * PS + (size_t)((N - m.j) % N) * N + (N - m.k) % N;
    for (int a = 0; a < 3; ++a) {
        const int x = 2 * a, y = 2 * a + 1;
        U[a * astride + at] = cpx<T>{(T)(ur[x] - ui[y]), (T)(ui[x] + ur[y])};
        PU[a * astride + at] = cpx<T>{(T)(pr[x] - pim[y]), (T)(pim[x] + pr[y])};
        if (!m.real) {
            U[a * astride + cjx] = cpx<T>{(T)(ur[x] + ui[y]), (T)(ur[y] - ui[x])};
            PU[a * astride + cjx] = cpx<T>{(T)(pr[x] + pim[y]), (T)(pr[y] - pim[x])};
        }
    }
}

}  // namespace psc
