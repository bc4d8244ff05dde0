// psc_adjoint.hpp -- discrete adjoint of the spectral lattice step.
//
// Reverse-mode derivative of evolve_VV_k_nopad (evolution.cpp:1131-1198) with
// the Friedmann coupling of kernel_a_k_nopad (evolution.cpp:1061-1090), on the
// packed spectra.  Conventions: for a real objective J and a complex array z,
// zbar = dJ/dRe z + i dJ/dIm z.  With r = F^H f / N^3 the packed real-space
// fields, the force F(f, a) = -a^{1+al} kIR^2 n^2 f - a^{3+al} F(V'(r)) has the
// adjoint  fbar += -a^{1+al} kIR^2 n^2 Fbar + F(Hess V(r) . F^H(-a^{3+al} Fbar)) / N^3.
//
// One backward substep s (end point e = s + 1) runs five passes:
//   P1 z-axis row pass: the add-adjoint and kick-adjoint updates of pibar,
//      fbar; Fbar = h pibar + h_next pibar_old (the first kick of substep s + 1
//      evaluates the force at the same point, so its adjoint h_next pibar_old,
//      pibar as that substep's sweep left it, is merged here);
//      ybar = -a'^{3+al} Fbar, z inverse.
//   P2 y-axis inverse (k_strided).
//   P3 x-axis inverse -> vbar = F^H ybar; q = Hess V(r') vbar + cV V'(r')
//      (the V part of the add adjoint); x forward.  Reductions for abar and
//      the coupling's adjoint.
//   P4 y-axis forward (k_strided).
//   P5 z-axis forward -> fbar += q_k / N^3; the drift adjoint
//      pibar_h = pibar + 2 h a_h^{al-3} fbar.
// The scalar recursion (abar, bbar, addbar, g2bar) runs in k_adj_scalars.
//
// With the GW-energy-fraction objective the six u_ij (three packed pairs U_q)
// carry adjoints ubar_q, pibar_u,q.  The GW kick pi_u += h a^{1+al} (-kIR^2 n^2 u
// + 2 S(f)) at the end point (merged with the next substep's first kick, as for
// the fields) gives ubar += -a^{1+al} kIR^2 n^2 Pib and Sbar = 2 a^{1+al} Pib with
// Pib = (h + h_next) pibar_u; the source adjoint reaches the fields through the
// gradients, fbar += sum_d (-i kIR n_d) F(gbar_d) / N^3 with
// gbar_d = sum_e w_de sbar_de G_e (w_dd = 2), in three more passes (P1g, P3g,
// P5g) that run beside P1, P3 and P5.
#pragma once
#include "psc_kernels.hpp"

namespace psc {

// Coefficients of one backward substep, written by k_adj_scalars.
struct AdjCoef {
    double h;    // w dt / 2
    double hn;   // w dt / 2 of the substep undone before (0 at the end of the run)
    double cs;   // 2 a'^{1+al}                         (GW source)
    double ck;   // addbar cA (al - 2) / (a'^6 N^6)    (K part of the add adjoint)
    double cg;   // addbar cA al kIR^2 / (a'^2 N^6)     (G part)
    double c1;   // a'^{1+al} kIR^2
    double c3;   // a'^{3+al}
    double cV;   // addbar cA (al + 1) / N^3            (V part)
    double cd;   // 2 h a_h^{al-3}                      (drift)
    int final_pass;  // t = 0: no substep to undo, only the first-kick force and add0
    int pad;
};

// Scalar adjoint state.
struct AdjState {
    double abar, bbar, addbar, g2bar;   // g2bar: the adjoint of the chosen coupling
    double addbar_used;   // addbar of the substep being undone (V part of its add adjoint)
    double a_e;           // forward a at that substep's end point
    double h_undo;        // w dt / 2 of that substep
    double ah_undo;       // its a_half
    int pending;          // a substep has been prepared and awaits its POST update
    int pad_i;
    double pad[7];
};

// P1
template <class G>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_adj_p1(const cpx<typename G::T>* __restrict__ fe, const cpx<typename G::T>* __restrict__ pe,
         cpx<typename G::T>* __restrict__ fbar, cpx<typename G::T>* __restrict__ pbar,
         cpx<typename G::T>* __restrict__ W, const AdjCoef* __restrict__ co,
         const cpx<typename G::T>* __restrict__ tw, double* __restrict__ part) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (line / N) * G::PS + (line % N) * N + t;
    const T h = (T)co->h, hn = (T)co->hn, ck = (T)co->ck, cg = (T)co->cg, c1 = (T)co->c1, c3 = (T)co->c3;
    const int px = signed_index((int)(line / N), N), py = signed_index((int)(line % N), N);
    const int pxy = px * px + py * py;
    double red[1] = {0.0};
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const size_t idx = base + G::TPL * m;
        const int pz = signed_index(t + G::TPL * m, N);
        const T n2 = (T)(pxy + pz * pz);
        const cpx<T> f = fe[idx], p = pe[idx], po = pbar[idx];
        cpx<T> fb = fbar[idx];
        const cpx<T> pb = {po.x + ck * p.x, po.y + ck * p.y};
        pbar[idx] = pb;
        const cpx<T> Fb = {h * pb.x + hn * po.x, h * pb.y + hn * po.y};
        fb.x += cg * n2 * f.x - c1 * n2 * Fb.x;
        fb.y += cg * n2 * f.y - c1 * n2 * Fb.y;
        fbar[idx] = fb;
        red[0] += (double)n2 * ((double)Fb.x * f.x + (double)Fb.y * f.y);
        e[m] = cpx<T>{-c3 * Fb.x, -c3 * Fb.y};
    }
    line_fft<T, N, typename G::Rad, +1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + G::TPL * m] = e[m];
    block_sum<1, G::THREADS_ROW>(red, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// P3
template <class G>
__global__ void __launch_bounds__((G::THREADS_COL))
k_adj_p3(cpx<typename G::T>* __restrict__ W, const cpx<typename G::T>* __restrict__ re,
         const MemberState* __restrict__ ms, const AdjCoef* __restrict__ co,
         const cpx<typename G::T>* __restrict__ tw, double* __restrict__ part, int coupling) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = G::PS;
    const size_t base = o * N + z + t * stride;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * stride];
    const Couplings<T> cp = couplings<T>(*ms);   // read before the transform hides their latency
    const T cV = (T)co->cV;
    line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, tw);
    double red[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const cpx<T> r = re[base + (size_t)G::TPL * m * stride];
        const T f[2] = {r.x, r.y};
        const cpx<T> vb = e[m];
        // V' and the Hessian (dvdf, d2Vdf2 with index fld + fld2, model.h:75-93)
        const T vpx = Model::dvdf<T>(0, f, cp), vpy = Model::dvdf<T>(1, f, cp);
        red[0] += (double)vb.x * vpx + (double)vb.y * vpy;
        red[1] += (double)vb.x * d2V_dfdcoupling<Model, T>(coupling, 0, f, cp) +
                  (double)vb.y * d2V_dfdcoupling<Model, T>(coupling, 1, f, cp);
        red[2] += (double)dV_dcoupling<Model, T>(coupling, f, cp);
        const T hpp = Model::d2Vdf2<T>(0, f, cp), hpc = Model::d2Vdf2<T>(1, f, cp), hcc = Model::d2Vdf2<T>(2, f, cp);
        e[m] = cpx<T>{hpp * vb.x + hpc * vb.y + cV * vpx, hpc * vb.x + hcc * vb.y + cV * vpy};
    }
    line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + (size_t)G::TPL * m * stride] = e[m];
    block_sum<3, G::THREADS_COL>(red, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) {
        double* dst = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 3;
        dst[0] = red[0];
        dst[1] = red[1];
        dst[2] = red[2];
    }
}

// P5: z-axis forward of W0, fbar += W0 / N^3, then the drift adjoint
// pbar += cd fbar; reduction sum Re(conj(fbar) pi_half).  GW: W0 also holds
// the x and y shares of the gradient adjoint (k_adj_p3g, k_adj_y_fwd_gw), W3
// its z share, which enters as -i kIR n_z FFT_z(W3) (Nyquist component zeroed,
// evolution.cpp:494-495), so that fbar is complete before the drift adjoint;
// then the drift adjoint of u: pibar_u += cd ubar, reduction
// sum Re(conj(ubar) pi_u half) into part_g.
template <class G, bool GW = false>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_adj_p5(const cpx<typename G::T>* __restrict__ W, cpx<typename G::T>* __restrict__ fbar,
         cpx<typename G::T>* __restrict__ pbar, const cpx<typename G::T>* __restrict__ ph,
         const AdjCoef* __restrict__ co,
         const cpx<typename G::T>* __restrict__ tw, double* __restrict__ part,
         size_t astride = 0, double kIR = 0.0, const cpx<typename G::T>* __restrict__ ubar = nullptr,
         cpx<typename G::T>* __restrict__ pubar = nullptr, const cpx<typename G::T>* __restrict__ puh = nullptr,
         double* __restrict__ part_g = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (line / N) * G::PS + (line % N) * N + t;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + G::TPL * m];
    line_fft<T, N, typename G::Rad, -1, typename G::LayRow>(e, t, c, lds, tw);
    if constexpr (GW) {
        cpx<T> g[G::E];
#pragma unroll
        for (int m = 0; m < G::E; ++m) g[m] = W[3 * astride + base + G::TPL * m];
        line_fft<T, N, typename G::Rad, -1, typename G::LayRow>(g, t, c, lds, tw);
#pragma unroll
        for (int m = 0; m < G::E; ++m) {
            const T kz = (T)(kIR * deriv_index(t + G::TPL * m, N));
            e[m].x += kz * g[m].y;   // (-i kz) (x + i y) = kz y - i kz x
            e[m].y -= kz * g[m].x;
        }
    }
    const T inv = (T)(1.0 / (double)G::N3);
    const T cd = (T)co->cd;
    const bool fin = co->final_pass != 0;
    double red[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const size_t idx = base + G::TPL * m;
        cpx<T> fb = fbar[idx];
        fb.x += inv * e[m].x;
        fb.y += inv * e[m].y;
        fbar[idx] = fb;
        if (!fin) {
            const cpx<T> p = ph[idx];
            red[0] += (double)fb.x * p.x + (double)fb.y * p.y;
            cpx<T> pb = pbar[idx];
            pb.x += cd * fb.x;
            pb.y += cd * fb.y;
            pbar[idx] = pb;
        }
    }
    if constexpr (GW) {
        if (!fin) {
#pragma unroll 1
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int m = 0; m < G::E; ++m) {
                    const size_t idx = (size_t)q * astride + base + G::TPL * m;
                    const cpx<T> ub = ubar[idx], pu = puh[idx];
                    cpx<T> pb = pubar[idx];
                    pb.x += cd * ub.x;
                    pb.y += cd * ub.y;
                    pubar[idx] = pb;
                    red[1] += (double)ub.x * pu.x + (double)ub.y * pu.y;
                }
        }
    }
    block_sum<2, G::THREADS_ROW>(red, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) {
        part[blockIdx.x] = red[0];
        if constexpr (GW) part_g[blockIdx.x] = red[1];
    }
}

// P1 (GW): Pib = (h + h_next) pibar_u; ubar += -a^{1+al} kIR^2 n^2 Pib; the
// source adjoint Sbar = 2 a^{1+al} Pib goes through the z-axis inverse into
// W[1..3]; reduction sum n^2 Re(conj(Pib) u_e) for abar.
template <class G>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_adj_p1g(const cpx<typename G::T>* __restrict__ pubar, cpx<typename G::T>* __restrict__ ubar,
          const cpx<typename G::T>* __restrict__ ue, cpx<typename G::T>* __restrict__ W, size_t astride,
          const AdjCoef* __restrict__ co, const cpx<typename G::T>* __restrict__ tw, double* __restrict__ part) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (line / N) * G::PS + (line % N) * N + t;
    const T hh = (T)(co->h + co->hn), c1 = (T)co->c1, cs = (T)co->cs;
    const int px = signed_index((int)(line / N), N), py = signed_index((int)(line % N), N);
    const int pxy = px * px + py * py;
    double red[1] = {0.0};
#pragma unroll 1
    for (int q = 0; q < 3; ++q) {
        cpx<T> e[G::E];
#pragma unroll
        for (int m = 0; m < G::E; ++m) {
            const size_t idx = (size_t)q * astride + base + G::TPL * m;
            const int pz = signed_index(t + G::TPL * m, N);
            const T n2 = (T)(pxy + pz * pz);
            const cpx<T> pb = pubar[idx], u = ue[idx];
            const cpx<T> Pib = {hh * pb.x, hh * pb.y};
            cpx<T> ub = ubar[idx];
            ub.x -= c1 * n2 * Pib.x;
            ub.y -= c1 * n2 * Pib.y;
            ubar[idx] = ub;
            red[0] += (double)n2 * ((double)Pib.x * u.x + (double)Pib.y * u.y);
            e[m] = cpx<T>{cs * Pib.x, cs * Pib.y};
        }
        line_fft<T, N, typename G::Rad, +1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
        for (int m = 0; m < G::E; ++m) W[(size_t)(q + 1) * astride + base + G::TPL * m] = e[m];
    }
    block_sum<1, G::THREADS_ROW>(red, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// P3 (GW): x-axis inverse of sbar (three packed pairs in W1..W3), the
// gradient adjoint gbar_d = sum_e w_de sbar_de G_e with the recorded
// real-space gradients G_e, x-axis forward of gbar; reduction sum sbar . s
// for abar.  Each gbar_d reaches fbar as -i kIR n_d FFT_z FFT_y (gbar_d) / N^3,
// the field adjoint W0 as FFT_z FFT_y (W0) / N^3 (k_adj_p5).  n_x is constant
// along y and z lines, so -i kIR n_x gbar_x is added into W0 here; the y
// forward adds the y share into W0 as well (k_adj_y_fwd_gw), and the z share
// goes on in W3.  Arguments: W = W1, W0 = the field adjoint array.
template <class G>
__global__ void __launch_bounds__((G::THREADS_COL))
k_adj_p3g(cpx<typename G::T>* __restrict__ W, cpx<typename G::T>* __restrict__ W0,
          const cpx<typename G::T>* __restrict__ ge, size_t astride, double kIR,
          const cpx<typename G::T>* __restrict__ tw, double* __restrict__ part) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = G::PS;
    const size_t base = o * N + z + t * stride;
    cpx<T> e[3][G::E];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int m = 0; m < G::E; ++m) e[q][m] = W[(size_t)q * astride + base + (size_t)G::TPL * m * stride];
        line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e[q], t, c, lds, tw);
    }
    double red[1] = {0.0};
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const size_t idx = base + (size_t)G::TPL * m * stride;
        const cpx<T> g0 = ge[idx], g1 = ge[astride + idx], g2 = ge[2 * astride + idx];
        const T s11 = e[0][m].x, s22 = e[0][m].y, s33 = e[1][m].x, s12 = e[1][m].y, s13 = e[2][m].x,
                s23 = e[2][m].y;
        const T p00 = g0.x * g0.x + g0.y * g0.y, p11 = g1.x * g1.x + g1.y * g1.y, p22 = g2.x * g2.x + g2.y * g2.y;
        const T p01 = g0.x * g1.x + g0.y * g1.y, p02 = g0.x * g2.x + g0.y * g2.y, p12 = g1.x * g2.x + g1.y * g2.y;
        red[0] += (double)s11 * p00 + (double)s22 * p11 + (double)s33 * p22 + (double)s12 * p01 +
                  (double)s13 * p02 + (double)s23 * p12;
        const T two = (T)2;
        e[0][m] = cpx<T>{two * s11 * g0.x + s12 * g1.x + s13 * g2.x, two * s11 * g0.y + s12 * g1.y + s13 * g2.y};
        e[1][m] = cpx<T>{two * s22 * g1.x + s12 * g0.x + s23 * g2.x, two * s22 * g1.y + s12 * g0.y + s23 * g2.y};
        e[2][m] = cpx<T>{two * s33 * g2.x + s13 * g0.x + s23 * g1.x, two * s33 * g2.y + s13 * g0.y + s23 * g1.y};
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(e[q], t, c, lds, tw);
        if (q == 0) {
#pragma unroll
            for (int m = 0; m < G::E; ++m) {   // W0 += (-i kx) (x + i y) = kx y - i kx x
                const size_t idx = base + (size_t)G::TPL * m * stride;
                const T kx = (T)(kIR * deriv_index(t + G::TPL * m, N));
                cpx<T> w0 = W0[idx];
                w0.x += kx * e[0][m].y;
                w0.y -= kx * e[0][m].x;
                W0[idx] = w0;
            }
        } else {
#pragma unroll
            for (int m = 0; m < G::E; ++m) W[(size_t)q * astride + base + (size_t)G::TPL * m * stride] = e[q][m];
        }
    }
    block_sum<1, G::THREADS_COL>(red, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// y forward of the GW backward pass: W0 <- FFT_y(W0) - i kIR n_y FFT_y(W2)
// (W0 holds the field adjoint and the x share of the gradient adjoint, W2 its
// y share), W3 (the z share) in place.  W1 is not read.
template <class G>
__global__ void __launch_bounds__((G::THREADS_COL))
k_adj_y_fwd_gw(cpx<typename G::T>* __restrict__ W, size_t astride, double kIR,
               const cpx<typename G::T>* __restrict__ tw) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t base = (size_t)blockIdx.y * G::PS + z + (size_t)t * N;
    cpx<T> e[G::E], g[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * N];
    line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) g[m] = W[2 * astride + base + (size_t)G::TPL * m * N];
    line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(g, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const T ky = (T)(kIR * deriv_index(t + G::TPL * m, N));
        e[m].x += ky * g[m].y;
        e[m].y -= ky * g[m].x;
        W[base + (size_t)G::TPL * m * N] = e[m];
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[3 * astride + base + (size_t)G::TPL * m * N];
    line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[3 * astride + base + (size_t)G::TPL * m * N] = e[m];
}


// ---------------------------------------------------------------------------
// GW-energy-fraction objective at the end of the run: Q = sum over the modes
// the reference visits (half spectrum kz <= N/2, one of each conjugate pair,
// weight 2, self-conjugate corners weight 1; output.cpp:800-1040) of the
// projected |pi_u(k)|^2, plus the unprojected k = 0 term (output.cpp:1032-1038),
// and its gradient scale * dQ / d(pi_u pairs) written into G (three packed
// arrays).  Every element of G belongs to exactly one visited mode (as k or
// as -k), so each thread writes its own elements.
struct zd {
    double x, y;
};
__device__ __forceinline__ zd zadd(zd a, zd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ zd zsub(zd a, zd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ zd zscale(double s, zd a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ zd zmul(zd a, zd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ zd zconj(zd a) { return {a.x, -a.y}; }
__device__ __forceinline__ double znorm(zd a) { return a.x * a.x + a.y * a.y; }

template <class G>
__global__ void __launch_bounds__(256)
k_gw_objective(const cpx<typename G::T>* __restrict__ PU, cpx<typename G::T>* __restrict__ Gout, int projection,
               double scale, double* __restrict__ part) {
    using T = typename G::T;
    constexpr int N = G::N, H = N / 2;
    constexpr size_t as = G::MS;
    const size_t li = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double q = 0.0;
    constexpr size_t total = (size_t)N * N * (H + 1);
    if (li < total) {
        const int k = (int)(li % (H + 1)), j = (int)((li / (H + 1)) % N), i = (int)(li / ((size_t)(H + 1) * N));
        double wgt = 0.0;
        bool dc = false;
        if (k > 0 && k < H) {
            wgt = 2.0;
        } else {
            const bool ie = i == 0 || i == H, je = j == 0 || j == H;
            if (ie && je) {
                if (i == 0 && j == 0 && k == 0) dc = true;
                else wgt = 1.0;
            } else if (ie && j >= 1 && j < H) {
                wgt = 2.0;
            } else if (je && i >= 1 && i < H) {
                wgt = 2.0;
            } else if (i >= 1 && i < H && ((j >= 1 && j < H) || (j > H && j < N))) {
                wgt = 2.0;
            }
        }
        if (wgt > 0.0 || dc) {
            const size_t kk = (size_t)i * G::PS + (size_t)j * N + k;
            const size_t mk = (size_t)((N - i) % N) * G::PS + (size_t)((N - j) % N) * N + (N - k) % N;
            // the six components 11, 22, 33, 12, 13, 23 from the pairs (11,22), (33,12), (13,23)
            zd u[6];
#pragma unroll
            for (int pq = 0; pq < 3; ++pq) {
                const cpx<T> a = PU[pq * as + kk], b = PU[pq * as + mk];
                const zd x = {(double)a.x, (double)a.y}, y = {(double)b.x, -(double)b.y};
                u[2 * pq] = zscale(0.5, zadd(x, y));
                const zd d = zsub(x, y);
                u[2 * pq + 1] = {0.5 * d.y, -0.5 * d.x};   // (x - y) / 2i
            }
            zd gc[6];
            if (dc) {
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const double w = c < 3 ? 1.0 : 2.0;   // diagonal once, off-diagonal twice
                    q += w * znorm(u[c]);
                    gc[c] = zscale(2.0 * w, u[c]);
                }
            } else {
                const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
                const double kv[3] = {(double)signed_index(i, N), (double)signed_index(j, N), (double)k};
                const double kn2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                double P[3][3];
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int y = 0; y < 3; ++y) P[x][y] = (x == y ? 1.0 : 0.0) - kv[x] * kv[y] / kn2;
                zd M[3][3];
#pragma unroll
                for (int c = 0; c < 6; ++c) M[ci[c]][cj[c]] = M[cj[c]][ci[c]] = u[c];
                zd PM[3][3];
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int y = 0; y < 3; ++y)
                        PM[x][y] = zadd(zadd(zscale(P[x][0], M[0][y]), zscale(P[x][1], M[1][y])), zscale(P[x][2], M[2][y]));
                const zd tr = zadd(zadd(PM[0][0], PM[1][1]), PM[2][2]);
                zd Gm[3][3];
                if (projection == 1) {   // TT: A = P M P - P tr(P M) / 2; gradient 2 (P A P - P tr(P A) / 2)
                    zd A[3][3];
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y) {
                            A[x][y] = zadd(zadd(zscale(P[0][y], PM[x][0]), zscale(P[1][y], PM[x][1])),
                                           zscale(P[2][y], PM[x][2]));
                            A[x][y] = zsub(A[x][y], zscale(0.5 * P[x][y], tr));
                            q += znorm(A[x][y]);
                        }
                    zd PA[3][3];
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y)
                            PA[x][y] = zadd(zadd(zscale(P[x][0], A[0][y]), zscale(P[x][1], A[1][y])), zscale(P[x][2], A[2][y]));
                    const zd trPA = zadd(zadd(PA[0][0], PA[1][1]), PA[2][2]);
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y) {
                            const zd pap = zadd(zadd(zscale(P[0][y], PA[x][0]), zscale(P[1][y], PA[x][1])),
                                                zscale(P[2][y], PA[x][2]));
                            Gm[x][y] = zscale(2.0, zsub(pap, zscale(0.5 * P[x][y], trPA)));
                        }
                } else {                 // one-sided P M: |P M|^2 - |tr P M|^2 / 2; gradient 2 P (P M) - tr(P M) P
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y) q += znorm(PM[x][y]);
                    q -= 0.5 * znorm(tr);
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y) {
                            const zd ppm = zadd(zadd(zscale(P[x][0], PM[0][y]), zscale(P[x][1], PM[1][y])),
                                                zscale(P[x][2], PM[2][y]));
                            Gm[x][y] = zsub(zscale(2.0, ppm), zscale(P[x][y], tr));
                        }
                }
                q *= wgt;
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    gc[c] = zscale(wgt, ci[c] == cj[c] ? Gm[ci[c]][ci[c]] : zadd(Gm[ci[c]][cj[c]], Gm[cj[c]][ci[c]]));
            }
            // chain rule of the unpacking: pair += g_even / 2 + i g_odd / 2 at k, conjugates at -k
#pragma unroll
            for (int pq = 0; pq < 3; ++pq) {
                const zd ge = gc[2 * pq], go = gc[2 * pq + 1];
                const zd at_k = zadd(zscale(0.5, ge), zmul({0.0, 0.5}, go));
                const zd at_mk = zadd(zscale(0.5, zconj(ge)), zmul({0.0, 0.5}, zconj(go)));
                if (kk == mk) {
                    const zd v = zadd(at_k, at_mk);
                    Gout[pq * as + kk] = cpx<T>{(T)(scale * v.x), (T)(scale * v.y)};
                } else {
                    Gout[pq * as + kk] = cpx<T>{(T)(scale * at_k.x), (T)(scale * at_k.y)};
                    Gout[pq * as + mk] = cpx<T>{(T)(scale * at_mk.x), (T)(scale * at_mk.y)};
                }
            }
        }
    }
    double r[1] = {q};
    __shared__ double sh[4];
    block_sum<1, 256>(r, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

// Partial sums of dV/d(coupling) over recorded real-space fields.
template <class G>
__global__ void __launch_bounds__(256)
k_vg_sums(const cpx<typename G::T>* __restrict__ r, const MemberState* __restrict__ ms, int coupling,
          double* __restrict__ part) {
    constexpr int N = G::N;
    const Couplings<double> cp(ms->par);
    double acc[1] = {0.0};
    for (size_t li = (size_t)blockIdx.x * blockDim.x + threadIdx.x; li < G::N3; li += (size_t)gridDim.x * blockDim.x) {
        const size_t idx = (li / ((size_t)N * N)) * G::PS + li % ((size_t)N * N);
        const double f[2] = {(double)r[idx].x, (double)r[idx].y};
        acc[0] += dV_dcoupling<Model, double>(coupling, f, cp);
    }
    __shared__ double sh[4];
    block_sum<1, 256>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Compensated single-block sums of the adjoint reductions: R0 (P1 partials),
// R1..R3 (P3 partials, width 3), R4 (P5 partials, when r5) and with GW
// R5 (P1g), R6 (P3g), R7 (P5g, when r5).
struct AdjPartials {
    const double* p1;
    const double* p3;
    const double* p5;
    const double* p1g;
    const double* p3g;
    const double* p5g;
};

static __device__ void adj_block_reduce(const AdjPartials& P, int n1, int n3, int n5, bool r5, bool gw, double* sh,
                                 double* R) {
    constexpr int NR = 8;
    Comp acc[NR];
    for (int i = threadIdx.x; i < n1; i += blockDim.x) {
        acc[0].add(P.p1[i]);
        if (gw) acc[5].add(P.p1g[i]);
    }
    for (int i = threadIdx.x; i < n3; i += blockDim.x) {
        acc[1].add(P.p3[3 * (size_t)i]);
        acc[2].add(P.p3[3 * (size_t)i + 1]);
        acc[3].add(P.p3[3 * (size_t)i + 2]);
        if (gw) acc[6].add(P.p3g[i]);
    }
    if (r5)
        for (int i = threadIdx.x; i < n5; i += blockDim.x) {
            acc[4].add(P.p5[i]);
            if (gw) acc[7].add(P.p5g[i]);
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double s2 = __shfl_xor(acc[q].s, off, 64), c2 = __shfl_xor(acc[q].c, off, 64);
            comp_merge(acc[q].s, acc[q].c, s2, c2);
        }
        if (lane == 0) {
            sh[(q * 4 + wave) * 2] = acc[q].s;
            sh[(q * 4 + wave) * 2 + 1] = acc[q].c;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < NR; ++q) {
            double ss = sh[(q * 4) * 2], cc = sh[(q * 4) * 2 + 1];
            for (int w = 1; w < 4; ++w) comp_merge(ss, cc, sh[(q * 4 + w) * 2], sh[(q * 4 + w) * 2 + 1]);
            R[q] = ss + cc;
        }
}

// Scalar adjoint recursion.  mode bit 1: POST for the substep just undone
// (uses the reductions of P1 / P3 / P5; skipped while nothing is pending);
// bit 2: PRE for the next substep to undo; bit 4: the t = 0 pass (PRE without
// a substep, POST without the drift and kick parts).  PRE takes the forward
// values at the substep's end point from the recorded scalar states: me is the
// state recorded after the substep (a at its end is me->a_start, or me->a for
// the last substep of a segment, where no next substep has begun), mh the one
// that began it (a_half); the t = 0 pass takes them from the op.
struct AdjScalarOp {
    int mode;
    int static_bg;
    int alpha;
    int end_of_segment;
    int gw;                      // GW adjoint passes ran (partials P1g, P3g, P5g)
    int pad3;
    double h_pre;                // w dt / 2 of the substep to undo next
    double a_e, K_e, G_e, V_e;   // t = 0 pass: forward values at t = 0
    double kIR, f2m, N3;
    int n1, n3, n5, pad2;        // partial counts of P1, P3, P5
};

static __global__ void __launch_bounds__(256)
k_adj_scalars(AdjState* __restrict__ st, AdjCoef* __restrict__ co, AdjPartials parts,
              const MemberState* __restrict__ me, const MemberState* __restrict__ mh, AdjScalarOp op) {
    __shared__ double sh[8 * 4 * 2];
    const int al = op.alpha;
    double R[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // states loaded up front (in flight during the reduction), written back once
    AdjState S = *st;
    const MemberState E = *me;
    const double ah_rec = mh->a_half;
    __syncthreads();   // every thread has read the state before thread 0 rewrites it
    if (!S.pending) op.mode &= ~1;
    if (op.mode & 1) adj_block_reduce(parts, op.n1, op.n3, op.n5, !(op.mode & 4), op.gw != 0, sh, R);
    if (threadIdx.x != 0) return;
    const double R1 = R[0], R2 = R[1], R3 = R[2], R4 = R[3], R5 = R[4] + R[7];
    if (op.mode & 1) {
        const double a = S.a_e;
        const double cA = 1.0 / 3.0 * op.f2m * ipow(a, 1 + 2 * al);
        // force adjoint: d/da of -a^{1+al} kIR^2 n^2 f and -a^{3+al} F(V'); g^2 via V' and add
        S.g2bar += R3 + S.addbar_used * cA * (al + 1) * R4 / op.N3;
        if (op.static_bg) {
            S.abar = S.bbar = S.addbar = 0.0;
        } else {
            S.abar += R1 * (-(1 + al) * ipow(a, al) * op.kIR * op.kIR) + (3 + al) / a * R2;
            // GW kick force a^{1+al} (-kIR^2 n^2 u + 2 S): d/da over the two inner products
            if (op.gw) S.abar += (1 + al) / a * (ipow(a, 1 + al) * (-op.kIR * op.kIR) * R[5] + R[6]);
            if (!(op.mode & 4)) {
                const double h = S.h_undo, ah = S.ah_undo;
                double ahbar = S.abar;                            // a' = ah + bh h
                double bhbar = S.bbar + h * S.abar;
                ahbar += 2.0 * h * (al - 3) * ipow(ah, al - 4) * R5;  // f' = f + 2h ah^{al-3} pi_h
                bhbar += h * ahbar;                               // ah = a + bh h
                S.abar = ahbar;
                S.bbar = bhbar;                                   // bh = b + h add
                S.addbar = h * bhbar;
            }
        }
    }
    if (op.mode & 2) {
        const bool fin = (op.mode & 4) != 0;
        if (!fin) {
            op.a_e = op.end_of_segment ? E.a : E.a_start;
            op.K_e = E.K;
            op.G_e = E.G;
            op.V_e = E.V;
        }
        const double a_h_pre = fin ? op.a_e : ah_rec;
        const double a = op.a_e;
        const double cA = 1.0 / 3.0 * op.f2m * ipow(a, 1 + 2 * al);
        const double h = fin ? 0.0 : op.h_pre;
        if (!fin) S.addbar += h * S.bbar;                        // b' = bh + h add'
        if (op.static_bg) S.addbar = 0.0;
        const double addbar = S.addbar;
        const double Q = (al - 2) * op.K_e + al * op.G_e + (al + 1) * op.V_e;
        if (!op.static_bg)                                        // add' = A(f', pi', a')
            S.abar += addbar * cA * ((1 + 2 * al) * Q / a - 6.0 * (al - 2) * op.K_e / a - 2.0 * al * op.G_e / a);
        S.addbar_used = addbar;
        S.a_e = a;
        AdjCoef C;
        C.hn = S.h_undo;   // the substep undone before this one (0 at the start of the sweep)
        S.h_undo = h;
        S.ah_undo = a_h_pre;
        C.h = h;
        C.ck = addbar * cA * (al - 2) / ipow(a, 6) / (op.N3 * op.N3);
        C.cg = addbar * cA * al * op.kIR * op.kIR / (a * a) / (op.N3 * op.N3);
        C.c1 = ipow(a, 1 + al) * op.kIR * op.kIR;
        C.c3 = ipow(a, 3 + al);
        C.cs = 2.0 * ipow(a, 1 + al);
        C.cV = addbar * cA * (al + 1) / op.N3;
        C.cd = fin ? 0.0 : 2.0 * h * ipow(a_h_pre, al - 3);
        C.final_pass = fin ? 1 : 0;
        C.pad = 0;
        *co = C;
    }
    S.pending = (op.mode & 2) ? 1 : 0;
    *st = S;
}

}  // namespace psc
