// psc_kernels.hpp -- the HIP kernels of the spectral lattice step.
//
// State per member (packed pair of real fields, full N^3 spectrum, natural
// order [kx][ky][kz]):  Pf = FFT(phi + i chi), Pp = FFT(pi_phi + i pi_chi).
// Every k-space operation of evolve_VV_k_nopad (evolution.cpp:1131-1198) has
// real coefficients that are the same for both fields, so it acts on the packed
// spectra directly; for nflds == 1 the imaginary channel stays zero.
//
// One substep of the fused pipeline (5 passes over HBM, W = work array):
//   A  z-axis inverse pass, drift fused:  Pf += c_drift Pp; W = IFFT_z(Pf)
//   B  y-axis inverse pass                 W = IFFT_y(W)
//   C  x-axis inverse + V'(phi,chi) + x-axis forward, <V> reduction
//   D  y-axis forward pass
//   E  z-axis forward pass, kick fused:   F = -a^{1+al} kIR^2 n^2 Pf - a^{3+al} W
//                                          Pp += c1 F (K, G reductions), Pp += c2 F
// The second kick of substep p and the first kick of substep p+1 use the same
// force (same f, same a) -- the reference recomputes it bit-identically at
// evolution.cpp:1144-1152 -- so one force evaluation serves both.
#pragma once
#include "psc_common.hpp"

namespace psc {

// ---------------------------------------------------------------------------
// A: contiguous-axis inverse pass; DRIFT fuses f += c_drift * pi (evolution.cpp:1164-1169).
template <typename T, int N, bool DRIFT>
__global__ void __launch_bounds__((Geo<T, N>::THREADS_ROW))
k_z_inverse(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ Pf, const cpx<T>* __restrict__ Pp,
            cpx<T>* __restrict__ W, const MemberState* __restrict__ ms,
            const cpx<T>* __restrict__ tw) {
    using G = Geo<T, N>;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const int member = blockIdx.y;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (size_t)member * G::N3 + line * N + t;
    cpx<T> e[G::E];
    if constexpr (DRIFT) {
        const T cd = (T)ms[member].c_drift;
#pragma unroll
        for (int m = 0; m < G::E; ++m) {
            cpx<T> f = Pf[base + G::TPL * m];
            const cpx<T> p = Pp[base + G::TPL * m];
            f.x += cd * p.x;
            f.y += cd * p.y;
            Pf[base + G::TPL * m] = f;
            e[m] = f;
        }
    } else {
#pragma unroll
        for (int m = 0; m < G::E; ++m) e[m] = src[base + G::TPL * m];
    }
    line_fft<T, N, +1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + G::TPL * m] = e[m];
}

// ---------------------------------------------------------------------------
// B / D: strided-axis pass in place.  AXIS 1 = y (stride N), 0 = x (stride N^2).
template <typename T, int N, int AXIS, int SIGN>
__global__ void __launch_bounds__((Geo<T, N>::THREADS_COL))
k_strided(cpx<T>* __restrict__ W, const cpx<T>* __restrict__ tw) {
    using G = Geo<T, N>;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = AXIS == 1 ? (size_t)N : (size_t)N * N;
    const size_t base = (size_t)member * G::N3 + (AXIS == 1 ? o * N * N : o * N) + z + t * stride;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * stride];
    line_fft<T, N, SIGN, typename G::LayCol>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + (size_t)G::TPL * m * stride] = e[m];
}

// ---------------------------------------------------------------------------
// C: x-axis inverse, pointwise V'(phi, chi) of model.h (dvdf, model.h:75-82;
// kernel_field_k_nopad evolution.cpp:1103-1109) and x-axis forward, with the
// lattice sum of V (kernel_a_k_nopad evolution.cpp:1077-1085) in FP64.
template <typename T, int N>
__global__ void __launch_bounds__((Geo<T, N>::THREADS_COL))
k_x_force(cpx<T>* __restrict__ W, const MemberState* __restrict__ ms, const cpx<T>* __restrict__ tw,
          double* __restrict__ partV) {
    using G = Geo<T, N>;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = (size_t)N * N;
    const size_t base = (size_t)member * G::N3 + o * N + z + t * stride;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * stride];
    line_fft<T, N, +1, typename G::LayCol>(e, t, c, lds, tw);
    const T inv = (T)(1.0 / (double)G::N3);
    const T m2 = (T)ms[member].m2, g2 = (T)ms[member].g2;
    const double m2d = ms[member].m2, g2d = ms[member].g2;
    double v[1] = {0.0};
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const T phi = e[m].x * inv, chi = e[m].y * inv;
        const T chi2 = chi * chi, phi2 = phi * phi;
        e[m].x = m2 * phi + g2 * phi * chi2;
        e[m].y = g2 * phi2 * chi;
        const double pd = phi, cdd = chi;
        v[0] += 0.5 * m2d * (pd * pd) + 0.5 * g2d * (pd * pd) * (cdd * cdd);
    }
    line_fft<T, N, -1, typename G::LayCol>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + (size_t)G::TPL * m * stride] = e[m];
    block_sum<1, G::THREADS_COL>(v, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0)
        partV[(size_t)member * G::BLOCKS_COL + blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

// ---------------------------------------------------------------------------
// E: z-axis forward pass with the kick fused (kernel_field_k_nopad
// evolution.cpp:1115-1127 and the kicks :1153-1158, :1184-1189).  W holds the
// unnormalised spectrum of V'.  c1/c2 are w dt / 2 of the two kicks applied
// with this force; REDUCE accumulates sum |pi|^2 after the first kick and
// sum n^2 |f|^2 (kinetic_energy_k / gradient_energy_k, evolution.cpp:923-1035).
template <typename T, int N, bool REDUCE>
__global__ void __launch_bounds__((Geo<T, N>::THREADS_ROW))
k_z_kick(const cpx<T>* __restrict__ W, const cpx<T>* __restrict__ Pf, cpx<T>* __restrict__ Pp,
         const MemberState* __restrict__ ms, const cpx<T>* __restrict__ tw,
         double* __restrict__ partKG, double c1d, double c2d, int use_a_start, StepConsts sc) {
    using G = Geo<T, N>;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const int member = blockIdx.y;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (size_t)member * G::N3 + line * N + t;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + G::TPL * m];
    line_fft<T, N, -1, typename G::LayRow>(e, t, c, lds, tw);
    const double a = use_a_start ? ms[member].a_start : ms[member].a;
    const T lapc = (T)(pow(a, 1 + sc.alpha) * sc.kIR * sc.kIR);
    const T vc = (T)pow(a, 3 + sc.alpha);
    const T c1 = (T)c1d, c2 = (T)c2d;
    const int ix = (int)(line / N), iy = (int)(line % N);
    const int px = signed_index(ix, N), py = signed_index(iy, N);
    const int pxy = px * px + py * py;
    double red[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const int kz = t + G::TPL * m;
        const int pz = signed_index(kz, N);
        const T n2 = (T)(pxy + pz * pz);
        const cpx<T> f = Pf[base + G::TPL * m];
        cpx<T> p = Pp[base + G::TPL * m];
        const T ln2 = lapc * n2;
        const cpx<T> F = {-ln2 * f.x - vc * e[m].x, -ln2 * f.y - vc * e[m].y};
        p.x += c1 * F.x;
        p.y += c1 * F.y;
        if constexpr (REDUCE) {
            const double px_ = p.x, py_ = p.y, fx = f.x, fy = f.y;
            red[0] += px_ * px_ + py_ * py_;
            red[1] += (double)(pxy + pz * pz) * (fx * fx + fy * fy);
        }
        p.x += c2 * F.x;
        p.y += c2 * F.y;
        Pp[base + G::TPL * m] = p;
    }
    if constexpr (REDUCE) {
        block_sum<2, G::THREADS_ROW>(red, reinterpret_cast<double*>(lds));
        if (threadIdx.x == 0) {
            double* dst = partKG + ((size_t)member * G::BLOCKS_ROW + blockIdx.x) * 2;
            dst[0] = red[0];
            dst[1] = red[1];
        }
    }
}

// ---------------------------------------------------------------------------
// Scalar update at a substep boundary, one block per member.  Finishes
// substep p (kernel_a_k_nopad evolution.cpp:1061-1090; b update :1193; end of
// step :1196-1197) and begins the next (b_half :1141, a_half :1161, drift
// coefficient :1164, a :1180).
struct ScalarOp {
    int finish;        // reduce partials and finish a substep
    int last_of_step;  // the finished substep closes a step
    int begin;         // begin a substep with weight w_begin
    int pad;
    double w_finish, w_begin;
    int nV, nKG;       // partial counts per member
};

__global__ void __launch_bounds__(256)
k_scalars(MemberState* __restrict__ ms, const double* __restrict__ partV,
          const double* __restrict__ partKG, ScalarOp op, StepConsts sc) {
    __shared__ double sh[3 * 2 * 256];
    const int member = blockIdx.x;
    MemberState& s = ms[member];
    if (op.finish) {
        Comp cv, ck, cg;
        for (int i = threadIdx.x; i < op.nV; i += blockDim.x) cv.add(partV[(size_t)member * op.nV + i]);
        for (int i = threadIdx.x; i < op.nKG; i += blockDim.x) {
            ck.add(partKG[((size_t)member * op.nKG + i) * 2]);
            cg.add(partKG[((size_t)member * op.nKG + i) * 2 + 1]);
        }
        sh[threadIdx.x * 6 + 0] = cv.s; sh[threadIdx.x * 6 + 1] = cv.c;
        sh[threadIdx.x * 6 + 2] = ck.s; sh[threadIdx.x * 6 + 3] = ck.c;
        sh[threadIdx.x * 6 + 4] = cg.s; sh[threadIdx.x * 6 + 5] = cg.c;
        __syncthreads();
        if (threadIdx.x == 0) {
            Comp V, K, Gs;
            for (int i = 0; i < (int)blockDim.x; ++i) {
                Comp x;
                x.s = sh[i * 6 + 0]; x.c = sh[i * 6 + 1]; V.merge(x);
                x.s = sh[i * 6 + 2]; x.c = sh[i * 6 + 3]; K.merge(x);
                x.s = sh[i * 6 + 4]; x.c = sh[i * 6 + 5]; Gs.merge(x);
            }
            const double a = s.a;
            const double Kn = 0.5 * K.value() / pow(a, 6) / (sc.N3 * sc.N3);
            const double Gn = sc.kIR * sc.kIR * Gs.value() / sc.N3 / 2.0 / (a * a) / sc.N3;
            const double Vn = V.value() / sc.N3;
            s.K = Kn;
            s.G = Gn;
            s.V = Vn;
            const int al = sc.alpha;
            double add = 1.0 / 3.0 * sc.f2m * pow(a, 1 + 2 * al) * ((al - 2) * Kn + al * Gn + (al + 1) * Vn);
            if (sc.expansion_static) add = 0.0;
            s.add = add;
            s.b = s.b_half + op.w_finish * sc.dt / 2.0 * add;
            if (op.last_of_step) {
                s.ad = sc.expansion_static ? 0.0 : s.b;
                s.t += sc.dt;
            }
        }
    }
    if (op.begin && threadIdx.x == 0) {
        const double w = op.w_begin, dt = sc.dt;
        s.b_half = s.b + w * dt / 2.0 * s.add;
        s.a_start = s.a;
        s.a_half = sc.expansion_static ? 1.0 : s.a + s.b_half * w * dt / 2.0;
        s.c_drift = w * dt * pow(s.a_half, sc.alpha - 3);
        s.a = sc.expansion_static ? 1.0 : s.a_half + s.b_half * w * dt / 2.0;
    }
}

// ---------------------------------------------------------------------------
// Diagnostics: x-axis inverse pass that normalises by 1/N^3, writes the real
// fields (phi, chi) back into W, and reduces sum phi, sum chi, sum phi^2,
// sum chi^2, sum V_0, sum V_1 (meansvars output.cpp:427-447, potential terms
// model.h:35-73).
template <typename T, int N>
__global__ void __launch_bounds__((Geo<T, N>::THREADS_COL))
k_x_inverse_diag(cpx<T>* __restrict__ W, const MemberState* __restrict__ ms,
                 const cpx<T>* __restrict__ tw, double* __restrict__ part) {
    using G = Geo<T, N>;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = (size_t)N * N;
    const size_t base = (size_t)member * G::N3 + o * N + z + t * stride;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * stride];
    line_fft<T, N, +1, typename G::LayCol>(e, t, c, lds, tw);
    const double inv = 1.0 / (double)G::N3;
    const double m2 = ms[member].m2, g2 = ms[member].g2;
    double r[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const double phi = (double)e[m].x * inv, chi = (double)e[m].y * inv;
        e[m].x = (T)phi;
        e[m].y = (T)chi;
        r[0] += phi;
        r[1] += chi;
        r[2] += phi * phi;
        r[3] += chi * chi;
        r[4] += 0.5 * m2 * (phi * phi);
        r[5] += 0.5 * g2 * (phi * phi) * (chi * chi);
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + (size_t)G::TPL * m * stride] = e[m];
    block_sum<6, G::THREADS_COL>(r, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) {
        double* dst = part + ((size_t)member * G::BLOCKS_COL + blockIdx.y * gridDim.x + blockIdx.x) * 6;
#pragma unroll
        for (int q = 0; q < 6; ++q) dst[q] = r[q];
    }
}

// Per-field kinetic and gradient sums over the full spectrum with the pair
// unpacked: phi_k = (P(k) + conj P(-k)) / 2, chi_k = (P(k) - conj P(-k)) / 2i.
// Sums over the full spectrum equal the reference's weighted half-spectrum
// sums (weight 2 for 0 < kz < N/2, evolution.cpp:949-966) for Hermitian data.
// part per block: K_phi, K_chi, G_phi, G_chi (raw sums, G without kIR^2).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_field_sums(const cpx<T>* __restrict__ Pf, const cpx<T>* __restrict__ Pp, double* __restrict__ part,
             int blocks_per_member) {
    constexpr size_t N3 = (size_t)N * N * N;
    const int member = blockIdx.y;
    const cpx<T>* f = Pf + (size_t)member * N3;
    const cpx<T>* p = Pp + (size_t)member * N3;
    double r[4] = {0, 0, 0, 0};
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < N3; idx += (size_t)blocks_per_member * 256) {
        const int i = (int)(idx / ((size_t)N * N)), j = (int)((idx / N) % N), k = (int)(idx % N);
        const size_t midx = ((size_t)((N - i) % N) * N + (N - j) % N) * N + (N - k) % N;
        const int px = signed_index(i, N), py = signed_index(j, N), pz = signed_index(k, N);
        const double n2 = (double)(px * px + py * py + pz * pz);
        {
            const double ar = f[idx].x, ai = f[idx].y, br = f[midx].x, bi = -(double)f[midx].y;
            const double phr = 0.5 * (ar + br), phi_ = 0.5 * (ai + bi);
            const double chr = 0.5 * (ai - bi), chi_ = -0.5 * (ar - br);
            r[2] += n2 * (phr * phr + phi_ * phi_);
            r[3] += n2 * (chr * chr + chi_ * chi_);
        }
        {
            const double ar = p[idx].x, ai = p[idx].y, br = p[midx].x, bi = -(double)p[midx].y;
            const double phr = 0.5 * (ar + br), phi_ = 0.5 * (ai + bi);
            const double chr = 0.5 * (ai - bi), chi_ = -0.5 * (ar - br);
            r[0] += phr * phr + phi_ * phi_;
            r[1] += chr * chr + chi_ * chi_;
        }
    }
    __shared__ double sh[4 * 4];
    block_sum<4, 256>(r, sh);
    if (threadIdx.x == 0) {
        double* dst = part + ((size_t)member * blocks_per_member + blockIdx.x) * 4;
        for (int q = 0; q < 4; ++q) dst[q] = r[q];
    }
}

// Binned spectra of one field (spectra(), output.cpp:626-705) over the half
// spectrum kz <= N/2 with weights 2 (0 < kz < N/2) / 1, bin
// l = floor(sqrt(n^2) + 0.5).  Per block partial histograms of
// (numpoints, sum |f_k|^2, sum |pi_k|^2, sum |a^{al-3} pi_k + (ad/a) f_k|^2).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_spectra(const cpx<T>* __restrict__ Pf, const cpx<T>* __restrict__ Pp,
          const MemberState* __restrict__ ms, int fld, int alpha, int nbins_total,
          double* __restrict__ part, int blocks_per_member) {
    extern __shared__ double hist[];  // nbins_total * 4
    constexpr size_t N3 = (size_t)N * N * N;
    constexpr int NH = N / 2 + 1;
    const int member = blockIdx.y;
    for (int i = threadIdx.x; i < nbins_total * 4; i += blockDim.x) hist[i] = 0.0;
    __syncthreads();
    const cpx<T>* f = Pf + (size_t)member * N3;
    const cpx<T>* p = Pp + (size_t)member * N3;
    const double a = ms[member].a, ad = ms[member].ad;
    const double aop = pow(a, alpha - 3), hub = ad / a;
    const size_t nk = (size_t)N * N * NH;
    for (size_t h = (size_t)blockIdx.x * blockDim.x + threadIdx.x; h < nk; h += (size_t)blocks_per_member * blockDim.x) {
        const int i = (int)(h / ((size_t)N * NH)), j = (int)((h / NH) % N), k = (int)(h % NH);
        const size_t idx = ((size_t)i * N + j) * N + k;
        const size_t midx = ((size_t)((N - i) % N) * N + (N - j) % N) * N + (N - k) % N;
        double fr, fi, pr, pi_;
        {
            const double ar = f[idx].x, ai = f[idx].y, br = f[midx].x, bi = -(double)f[midx].y;
            if (fld == 0) { fr = 0.5 * (ar + br); fi = 0.5 * (ai + bi); }
            else { fr = 0.5 * (ai - bi); fi = -0.5 * (ar - br); }
        }
        {
            const double ar = p[idx].x, ai = p[idx].y, br = p[midx].x, bi = -(double)p[midx].y;
            if (fld == 0) { pr = 0.5 * (ar + br); pi_ = 0.5 * (ai + bi); }
            else { pr = 0.5 * (ai - bi); pi_ = -0.5 * (ar - br); }
        }
        const int px = signed_index(i, N), py = signed_index(j, N);
        const int n2 = px * px + py * py + k * k;
        const int l = (int)floor(sqrt((double)n2) + 0.5);
        const double wgt = (k == 0 || k == N / 2) ? 1.0 : 2.0;
        const double fp2 = fr * fr + fi * fi, pp2 = pr * pr + pi_ * pi_;
        const double nr = aop * pr + hub * fr, ni = aop * pi_ + hub * fi;
        atomicAdd(&hist[l * 4 + 0], wgt);
        atomicAdd(&hist[l * 4 + 1], wgt * fp2);
        atomicAdd(&hist[l * 4 + 2], wgt * pp2);
        atomicAdd(&hist[l * 4 + 3], wgt * (nr * nr + ni * ni));
    }
    __syncthreads();
    double* dst = part + ((size_t)member * blocks_per_member + blockIdx.x) * nbins_total * 4;
    for (int i = threadIdx.x; i < nbins_total * 4; i += blockDim.x) dst[i] = hist[i];
}

// Fixed-order sum of per-block partial vectors: out[q] = sum_b part[b][q].
__global__ void k_sum_partials(const double* __restrict__ part, int nblocks, int width,
                               double* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= width) return;
    Comp acc;
    for (int b = 0; b < nblocks; ++b) acc.add(part[(size_t)b * width + q]);
    out[q] = acc.value();
}

// Packed-pair conversion FP64 -> T on the device (load path).
template <typename T>
__global__ void k_convert_in(const double* __restrict__ src, cpx<T>* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = cpx<T>{(T)src[2 * i], (T)src[2 * i + 1]};
}
template <typename T>
__global__ void k_convert_out(const cpx<T>* __restrict__ src, double* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        dst[2 * i] = (double)src[i].x;
        dst[2 * i + 1] = (double)src[i].y;
    }
}

}  // namespace psc
