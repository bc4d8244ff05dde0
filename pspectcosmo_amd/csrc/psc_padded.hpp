// psc_padded.hpp -- the dealiased stepper (evolve_VV_k_pad with padding,
// extract_padding, kernel_field_k_pad, kernel_a_k_pad: evolution.cpp:1271-1759)
// on the device, with N_pad = 2N or 3N/2 (any padded length that the
// Stockham transforms take, 2^k or 3 * 2^k).
//
// The N-grid state is the packed full spectrum P = phi_k + i chi_k (as for
// the unpadded stepper) plus, for the kz = 0 and kz = N/2 planes, a packed
// "ghost" Gamma = A_phi + i A_chi of the anti-Hermitian parts that the
// reference's half spectra acquire there: extract_padding reads the +N/2
// copy of every Nyquist index, which in those planes is not the conjugate of
// its mirror.  The ghost never reaches the N-grid real fields (c2r drops it),
// but it enters the padded field through the kz = N/2 split, the kick, and
// the half-spectrum sums of K and G, so it is carried to reproduce the
// reference's half spectra exactly.
//
// Per force evaluation, with X_p the padded array:
//   k_drift_pad     f += c_drift pi (N grid and ghosts)
//   k_pad_z         padding (Nyquist indices split into +-N/2 halves, the
//                   kz = N/2 halves (P +- Gamma) / 2), scaled by (N_pad/N)^3 so
//                   that the N_pad-grid transform's 1/N_pad^3 yields the
//                   reference's c2r / N^3; z-axis inverse transform
//   k_strided y+    y-axis inverse on the padded x planes |x| <= N/2
//   k_x_force       x-axis inverse (rows |x| > N/2 read as zero), V' residual
//                   and sum of V, x-axis forward (rows |x| <= N/2 stored)
//   k_strided y-    y-axis forward on the planes |x| <= N/2
//   k_unpad_z       z-axis forward of the lines with |x|, |y| <= N/2, keeping
//                   the modes |z| <= N/2 in R[(N+1)^3]
//   k_kick_pad      extract_padding from R, the spectral force and both kicks,
//                   the ghost kick, sums of |pi|^2 and n^2 |f|^2
#pragma once

#include "psc_kernels.hpp"

namespace psc {

__device__ __forceinline__ int pad_signed(int p, int NP) { return p < NP / 2 ? p : p - NP; }

// (N_pad / N)^3, the factor between the padded and the unpadded transform
// normalisations (extract_padding's normalize_factor, evolution.cpp:1602).
template <int NP, int N>
constexpr double pad_volume_ratio() {
    return ((double)NP / N) * ((double)NP / N) * ((double)NP / N);
}

// The padded kernels take the padded shape GP and the unpadded lattice N:
// every mode |n_d| <= N/2 has its own padded index (N_pad > N), and a block of
// L row lines lies within one padded x plane (L divides N_pad) or covers whole
// planes (N_pad divides L).
template <class GP, int N>
constexpr bool pad_pair_ok() {
    return GP::N > N && N % 2 == 0 && (GP::N % GP::L == 0 || GP::L % GP::N == 0);
}

// N-grid index (0..N-1) of a signed index s in [-N/2, N/2] (both ends -> N/2).
__device__ __forceinline__ int pad_wrap(int s, int N) { return s == -N / 2 ? N / 2 : (s + N) % N; }

// Drift of the N-grid state and of both ghost planes (evolution.cpp:1716-1720).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_drift_pad(cpx<T>* __restrict__ Pf, const cpx<T>* __restrict__ Pp, cpx<T>* __restrict__ Gf,
            const cpx<T>* __restrict__ Gp, MemberState* __restrict__ ms) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * N * N) return;
    const int x = (int)(e / ((size_t)N * N)), y = (int)((e / N) % N), z = (int)(e % N);
    const size_t idx = (size_t)x * plane_stride(N) + (size_t)y * N + z;
    const T cd = (T)ms[0].c_drift;
    cpx<T> f = Pf[idx];
    const cpx<T> p = Pp[idx];
    if (sizeof(T) == 4 && e == 0) {
        dc_drift(f, p, ms[0]);   // the homogeneous mode, compensated
    } else {
        f.x += cd * p.x;
        f.y += cd * p.y;
    }
    Pf[idx] = f;
    if (z == 0 || z == N / 2) {
        const size_t gi = (z == 0 ? 0 : (size_t)N * N) + (size_t)x * N + y;
        cpx<T> g = Gf[gi];
        const cpx<T> q = Gp[gi];
        g.x += cd * q.x;
        g.y += cd * q.y;
        Gf[gi] = g;
    }
}

// Padding and the z-axis inverse transform on the padded grid.  The k = 0
// mode goes to dcp (times (N_pad/N)^3) and is left out of the transform, as
// in the unpadded passes.  Only the padded lines that hold N-grid modes,
// (sx, sy) in [-N/2, N/2]^2, are transformed and stored: the launch covers
// exactly those (N + 1)^2 lines, L per block (pad_z_blocks), line li at
// sx = li / (N + 1) - N/2, sy = li % (N + 1) - N/2; the last block's lines
// past the end transform zeros (the transform's barriers need every thread)
// and store nothing.  The padded x planes |x| > N/2 and the rows |y| > N/2
// of the kept planes are never written: the y passes skip them and read them
// as zero.  (A launch over every padded line, most of whose blocks left at
// once, measured 140 us against 116 us at N = 256, N_pad = 512.)  STREAM
// stores the padded lines with non-temporal stores (88 us there): they are
// read once, by the y inverse, and exceed the memory-side cache.
template <class GP, int N>
constexpr unsigned pad_z_blocks() {
    return (unsigned)(((size_t)(N + 1) * (N + 1) + GP::L - 1) / GP::L);
}
template <class GP, int N, bool STREAM = false>
__global__ void __launch_bounds__((GP::THREADS_ROW))
k_pad_z(const cpx<typename GP::T>* __restrict__ Pf, const cpx<typename GP::T>* __restrict__ Gf,
        cpx<typename GP::T>* __restrict__ Wp, cpx<typename GP::T>* __restrict__ dcp,
        const cpx<typename GP::T>* __restrict__ tw) {
    using T = typename GP::T;
    constexpr int NP = GP::N, H = N / 2;
    static_assert(pad_pair_ok<GP, N>(), "padded shape and lattice do not pair");
    constexpr size_t PSN = plane_stride(N);
    __shared__ cpx<T> lds[GP::LayRow::size(GP::L)];
    const int t = threadIdx.x % GP::TPL, c = threadIdx.x / GP::TPL;
    const int li = (int)blockIdx.x * GP::L + c;
    const bool my = li < (N + 1) * (N + 1);
    const int sx = my ? li / (N + 1) - H : 0, sy = my ? li % (N + 1) - H : 0;
    const int xp = pad_wrap(sx, NP), yp = pad_wrap(sy, NP);
    const size_t base = (size_t)xp * GP::PS + (size_t)yp * NP + t;
    const int x = pad_wrap(sx, N), y = pad_wrap(sy, N);
    const T w = (T)(pad_volume_ratio<NP, N>() * (sx == H || sx == -H ? 0.5 : 1.0) *
                    (sy == H || sy == -H ? 0.5 : 1.0));
    cpx<T> e[GP::E];
#pragma unroll
    for (int m = 0; m < GP::E; ++m) {
        const int sz = pad_signed(t + GP::TPL * m, NP);
        cpx<T> v = {(T)0, (T)0};
        if (my && sz <= H && sz >= -H) {
            const int z = pad_wrap(sz, N);
            v = Pf[(size_t)x * PSN + (size_t)y * N + z];
            if (sz == H || sz == -H) {   // kz = N/2: (P +- Gamma) / 2
                const cpx<T> g = Gf[(size_t)N * N + (size_t)x * N + y];
                const T s = sz > 0 ? (T)1 : (T)-1;
                v.x = (T)0.5 * (v.x + s * g.x);
                v.y = (T)0.5 * (v.y + s * g.y);
            }
            v.x *= w;
            v.y *= w;
            if (x == 0 && y == 0 && sz == 0) {
                dcp[0] = v;
                v = cpx<T>{(T)0, (T)0};
            }
        }
        e[m] = v;
    }
    line_fft<T, NP, typename GP::Rad, +1, typename GP::LayRow>(e, t, c, lds, tw);
    if (!my) return;
#pragma unroll
    for (int m = 0; m < GP::E; ++m) {
        if constexpr (STREAM) store_stream(Wp + base + GP::TPL * m, e[m]);
        else Wp[base + GP::TPL * m] = e[m];
    }
}

// z-axis forward transform of the padded lines that hold N-grid modes, the
// modes |z| <= N/2 stored in R at ((sx + H)(N + 1) + sy + H)(N + 1) + sz + H.
// Launched over the (N + 1)^2 kept lines as k_pad_z (pad_z_blocks); STREAM
// reads the padded lines with non-temporal loads (read once).
template <class GP, int N, bool STREAM = false>
__global__ void __launch_bounds__((GP::THREADS_ROW))
k_unpad_z(const cpx<typename GP::T>* __restrict__ Wp, cpx<typename GP::T>* __restrict__ R,
          const cpx<typename GP::T>* __restrict__ tw) {
    using T = typename GP::T;
    constexpr int NP = GP::N, H = N / 2;
    static_assert(pad_pair_ok<GP, N>(), "padded shape and lattice do not pair");
    __shared__ cpx<T> lds[GP::LayRow::size(GP::L)];
    const int t = threadIdx.x % GP::TPL, c = threadIdx.x / GP::TPL;
    const int li = (int)blockIdx.x * GP::L + c;
    const bool my = li < (N + 1) * (N + 1);
    const int sx = my ? li / (N + 1) - H : 0, sy = my ? li % (N + 1) - H : 0;
    const size_t base = (size_t)pad_wrap(sx, NP) * GP::PS + (size_t)pad_wrap(sy, NP) * NP + t;
    cpx<T> e[GP::E];
#pragma unroll
    for (int m = 0; m < GP::E; ++m)   // lines past the end transform the (valid) line 0
        e[m] = STREAM ? load_stream(Wp + base + GP::TPL * m) : Wp[base + GP::TPL * m];
    line_fft<T, NP, typename GP::Rad, -1, typename GP::LayRow>(e, t, c, lds, tw);
    if (!my) return;
    const size_t row = ((size_t)(sx + H) * (N + 1) + (sy + H)) * (N + 1);
#pragma unroll
    for (int m = 0; m < GP::E; ++m) {
        const int sz = pad_signed(t + GP::TPL * m, NP);
        if (sz <= H && sz >= -H) R[row + (sz + H)] = e[m];
    }
}

// extract_padding and the kick (kernel_field_k_pad evolution.cpp:1665-1672,
// kicks :1704-1709, :1737-1742).  For the N-grid mode n with c = 2 per index
// equal to N/2 and r = (N_pad/N)^3: outside the kz = 0, N/2 planes
// E = c R(n with its Nyquist x, y indices at sign(kz) N/2) / r; in those planes
// E = c/2 (R(n+) + R(n-)) / r and the ghost source c/2 (R(n+) - R(n-)) / r,
// n+- carrying every Nyquist index at +-N/2.  The residual's k = 0 mode gets N^3 V'(mean) back.  Each
// thread handles 8 modes; REDUCE sums |pi|^2 after the first kick and
// n^2 |f|^2, ghosts included (the half-spectrum sums of the reference).
template <class GP, int N, bool REDUCE, bool SYNC_OUT = false, bool STREAM = false>
__global__ void __launch_bounds__(256)
k_kick_pad(const cpx<typename GP::T>* __restrict__ R, const cpx<typename GP::T>* __restrict__ Pf,
           cpx<typename GP::T>* __restrict__ Pp, const cpx<typename GP::T>* __restrict__ Gf,
           cpx<typename GP::T>* __restrict__ Gp, MemberState* __restrict__ ms,
           double* __restrict__ partKG, KickArgs ka, StepConsts sc,
           cpx<typename GP::T>* __restrict__ Ps = nullptr, cpx<typename GP::T>* __restrict__ Gps = nullptr) {
    using T = typename GP::T;
    constexpr int NP = GP::N, H = N / 2;
    static_assert(pad_pair_ok<GP, N>(), "padded shape and lattice do not pair");
    constexpr size_t PSN = plane_stride(N);
    __shared__ double red_lds[2 * 256 / 64 + 8];
    if (ka.prologue && ms[0].ahead) return;   // the call's first kick was applied as a look-ahead
    const double a = ka.prologue ? ms[0].a_start : ms[0].a;
    const T lapc = (T)(pow(a, 1 + sc.alpha) * sc.kIR * sc.kIR);
    const double vcd = pow(a, 3 + sc.alpha);
    const T vc = (T)vcd;
    double c1d, c2d;
    kick_coefs(ka, ms[0], c1d, c2d);
    const T c1 = (T)c1d, c2 = (T)c2d;
    const T invr = (T)(1.0 / pad_volume_ratio<NP, N>());
    auto ridx = [](int ax, int ay, int az) {
        return ((size_t)(ax + H) * (N + 1) + (ay + H)) * (N + 1) + (az + H);
    };
    double red[2] = {0.0, 0.0};
    // Two batches of 4 modes, each loading all its inputs before the kicks,
    // so that a thread has a batch's loads in flight at once (one mode at a
    // time measured 145 us, batches of 4 103 us, of 8 131 us (205 VGPRs), at
    // N = 256, N_pad = 512; bit-identical).  STREAM: R is read once, with
    // non-temporal loads (99 us).
    constexpr int BATCH = 4;
    auto ldR = [&](size_t i) { return STREAM ? load_stream(R + i) : R[i]; };
#pragma unroll 1
    for (int qb = 0; qb < 8; qb += BATCH) {
        cpx<T> Ra[BATCH], Rb[BATCH], Fq[BATCH], Pq[BATCH], Gfq[BATCH], Gpq[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
            const size_t e = ((size_t)blockIdx.x * 8 + qb + j) * 256 + threadIdx.x;
            const int x = (int)(e / ((size_t)N * N)), y = (int)((e / N) % N), z = (int)(e % N);
            const int sx = x <= H ? x : x - N, sy = y <= H ? y : y - N, sz = z <= H ? z : z - N;
            const bool X = x == H, Y = y == H, Z = z == H;
            if (!(z == 0 || Z)) {
                const int s = sz > 0 ? H : -H;
                Ra[j] = ldR(ridx(X ? s : sx, Y ? s : sy, sz));
            } else {
                Ra[j] = ldR(ridx(X ? H : sx, Y ? H : sy, Z ? H : 0));
                Rb[j] = ldR(ridx(X ? -H : sx, Y ? -H : sy, Z ? -H : 0));
                const size_t gi = (z == 0 ? 0 : (size_t)N * N) + (size_t)x * N + y;
                Gfq[j] = Gf[gi];
                Gpq[j] = Gp[gi];
            }
            const size_t idx = (size_t)x * PSN + (size_t)y * N + z;
            Fq[j] = Pf[idx];
            Pq[j] = Pp[idx];
        }
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
            const size_t e = ((size_t)blockIdx.x * 8 + qb + j) * 256 + threadIdx.x;
            const int x = (int)(e / ((size_t)N * N)), y = (int)((e / N) % N), z = (int)(e % N);
            const int sx = x <= H ? x : x - N, sy = y <= H ? y : y - N, sz = z <= H ? z : z - N;
            const bool X = x == H, Y = y == H, Z = z == H;
            const T cc = (T)((X ? 2 : 1) * (Y ? 2 : 1) * (Z ? 2 : 1));
            const bool plane = z == 0 || Z;
            cpx<T> E, GE = {(T)0, (T)0};
            if (!plane) {
                const cpx<T> r = Ra[j];
                E = cpx<T>{cc * r.x * invr, cc * r.y * invr};
            } else {
                const cpx<T> rp = Ra[j], rm = Rb[j];
                const T h = (T)0.5 * cc * invr;
                E = cpx<T>{h * (rp.x + rm.x), h * (rp.y + rm.y)};
                GE = cpx<T>{h * (rp.x - rm.x), h * (rp.y - rm.y)};
            }
            const size_t idx = (size_t)x * PSN + (size_t)y * N + z;
            const cpx<T> f = Fq[j];
            cpx<T> p = Pq[j];
            const int n2i = sx * sx + sy * sy + sz * sz;
            kick_mode<T, false, SYNC_OUT, REDUCE>(E, f, p, n2i, e == 0, ms[0], (double)N * N * N, lapc, vc, vcd,
                                                  c1d, c2d, red, nullptr, SYNC_OUT ? Ps + idx : nullptr);
            Pp[idx] = p;
            const T ln2 = lapc * (T)n2i;
            if (plane) {
                const size_t gi = (z == 0 ? 0 : (size_t)N * N) + (size_t)x * N + y;
                const cpx<T> gf = Gfq[j];
                cpx<T> gp = Gpq[j];
                const cpx<T> FG = {-ln2 * gf.x - vc * GE.x, -ln2 * gf.y - vc * GE.y};
                gp.x += c1 * FG.x;
                gp.y += c1 * FG.y;
                if constexpr (SYNC_OUT) Gps[gi] = gp;
                if constexpr (REDUCE) {
                    const double px_ = gp.x, py_ = gp.y, fx = gf.x, fy = gf.y;
                    red[0] += px_ * px_ + py_ * py_;
                    red[1] += (double)n2i * (fx * fx + fy * fy);
                }
                gp.x += c2 * FG.x;
                gp.y += c2 * FG.y;
                Gp[gi] = gp;
            }
        }
    }
    if constexpr (REDUCE) {
        block_sum<2, 256>(red, red_lds);
        if (threadIdx.x == 0) {
            partKG[(size_t)blockIdx.x * 2] = red[0];
            partKG[(size_t)blockIdx.x * 2 + 1] = red[1];
        }
    }
}

}  // namespace psc
