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
// Block `lin` of `total` in reverse order: the dispatcher deals consecutive
// blocks to the eight XCDs in turn, so groups of eight are reversed and the
// position inside a group kept; each block then lands on the XCD whose L2 saw
// the same lines in a pass of the same mapping run forward.
__device__ __forceinline__ unsigned reversed_block(unsigned lin, unsigned total) {
    return total % 8 == 0 ? total - 8 - (lin & ~7u) + (lin & 7u) : total - 1 - lin;
}

// ---------------------------------------------------------------------------
// A: contiguous-axis inverse pass; DRIFT fuses f += c_drift * pi (evolution.cpp:1164-1169).
// REV: lines in reverse order, as in k_strided.
template <class G, bool DRIFT, bool REC = false, bool REV = false, bool STREAM = false>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_z_inverse(const cpx<typename G::T>* __restrict__ src, cpx<typename G::T>* __restrict__ Pf, const cpx<typename G::T>* __restrict__ Pp,
            cpx<typename G::T>* __restrict__ W, MemberState* __restrict__ ms,
            const cpx<typename G::T>* __restrict__ tw, cpx<typename G::T>* __restrict__ rec_f = nullptr,
            cpx<typename G::T>* __restrict__ rec_ph = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    unsigned lin = blockIdx.x + gridDim.x * blockIdx.y;
    if constexpr (REV) lin = reversed_block(lin, gridDim.x * gridDim.y);
    const int member = (int)(lin / gridDim.x);
    const size_t line = (size_t)(lin % gridDim.x) * G::L + c;
    const size_t base = (size_t)member * G::MS + (line / N) * G::PS + (line % N) * N + t;
    cpx<T> e[G::E];
    if constexpr (DRIFT) {
        const T cd = (T)ms[member].c_drift;
#pragma unroll
        for (int m = 0; m < G::E; ++m) {
            // f and pi are read next by the z kick, three passes later: with
            // STREAM (arrays larger than the memory-side cache holds for
            // three passes) they do not allocate, so that the caches keep W
            // for the y inverse
            cpx<T> f, p;
            if constexpr (STREAM) {
                f = load_stream(Pf + base + G::TPL * m);
                p = load_stream(Pp + base + G::TPL * m);
            } else {
                f = Pf[base + G::TPL * m];
                p = Pp[base + G::TPL * m];
            }
            if constexpr (REC) {  // adjoint recording: substep start f and pi_half
                rec_f[base + G::TPL * m] = f;
                rec_ph[base + G::TPL * m] = p;
            }
            if (sizeof(T) == 4 && m == 0 && line == 0 && t == 0) {
                dc_drift(f, p, ms[member]);   // the homogeneous mode, compensated
            } else {
                f.x += cd * p.x;
                f.y += cd * p.y;
            }
            if constexpr (STREAM)
                store_stream(Pf + base + G::TPL * m, f);
            else
                Pf[base + G::TPL * m] = f;
            e[m] = f;
        }
    } else {
#pragma unroll
        for (int m = 0; m < G::E; ++m) e[m] = src[base + G::TPL * m];
    }
    // The k = 0 mode (the homogeneous field, N^3 times larger than the
    // fluctuations) stays out of the transforms: it is added back as an exact
    // constant in the x pass, so FP32 rounding of the transform scales with the
    // fluctuations, not with the mean.
    if (line == 0 && t == 0) e[0] = cpx<T>{(T)0, (T)0};
    line_fft<T, N, typename G::Rad, +1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + G::TPL * m] = e[m];
}

// Plain z-axis forward pass W -> dst (the last axis of a forward 3D transform
// of real-space data, used when a state is loaded from real space).
template <class G>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_z_forward(const cpx<typename G::T>* __restrict__ W, cpx<typename G::T>* __restrict__ dst,
            const cpx<typename G::T>* __restrict__ tw) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const int member = blockIdx.y;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (size_t)member * G::MS + (line / N) * G::PS + (line % N) * N + t;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + G::TPL * m];
    line_fft<T, N, typename G::Rad, -1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) dst[base + G::TPL * m] = e[m];
}

// ---------------------------------------------------------------------------
// B / D: strided-axis pass in place.  AXIS 1 = y (stride N), 0 = x (stride N^2).
// MAPPED (y passes of the dealiased stepper on its padded grid): blockIdx.y
// runs over the 2 MH + 1 x planes |x| <= MH that hold modes (MH = half the
// unpadded lattice: N/4 for N_pad = 2N, N/3 for N_pad = 3N/2); the others stay
// untouched.  Only the rows |y| <= MH of a spectrum hold modes: the inverse
// reads the others as zero (the padding pass does not write them) and the
// forward does not store them (the extraction pass does not read them).  REV: the blocks take the lattice from its far end (last member,
// last plane first, in the order of reversed_block), so that a pass starts on
// the lines the previous pass wrote last, which are still in the 256 MB
// memory-side cache.  STREAM: non-temporal loads and stores (the dealiased
// stepper's padded array, read once per write and beyond that cache:
// 187 -> 173 us inverse, 193 -> 185 us forward at N = 256, N_pad = 512).
template <class G, int AXIS, int SIGN, bool MAPPED = false, bool REV = false, int MH = G::N / 4, bool STREAM = false>
__global__ void __launch_bounds__((G::THREADS_COL))
k_strided(cpx<typename G::T>* __restrict__ W, const cpx<typename G::T>* __restrict__ tw) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    unsigned bx = blockIdx.x, by = blockIdx.y;
    int member = blockIdx.z;
    if constexpr (REV) {
        const unsigned lin = reversed_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                            gridDim.x * gridDim.y * gridDim.z);
        bx = lin % gridDim.x;
        by = (lin / gridDim.x) % gridDim.y;
        member = (int)(lin / (gridDim.x * gridDim.y));
    }
    const size_t z = (size_t)bx * G::TC + c;
    const size_t o = MAPPED ? (size_t)((by + N - MH) % N) : (size_t)by;
    constexpr size_t stride = AXIS == 1 ? (size_t)N : G::PS;
    const size_t base = (size_t)member * G::MS + (AXIS == 1 ? o * G::PS : o * N) + z + t * stride;
    static_assert(!MAPPED || AXIS == 1, "mapped passes run along y");
    auto kept = [&](int m) {   // spectrum row y = t + TPL m holds modes
        const int y = t + G::TPL * m;
        return !MAPPED || y <= MH || y >= N - MH;
    };
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m)
        e[m] = (SIGN < 0 || kept(m)) ? (STREAM ? load_stream(W + base + (size_t)G::TPL * m * stride)
                                               : W[base + (size_t)G::TPL * m * stride])
                                     : cpx<T>{(T)0, (T)0};
    line_fft<T, N, typename G::Rad, SIGN, typename G::LayCol>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m)
        if (SIGN > 0 || kept(m)) {
            if constexpr (STREAM) store_stream(W + base + (size_t)G::TPL * m * stride, e[m]);
            else W[base + (size_t)G::TPL * m * stride] = e[m];
        }
}

// The pointwise step of the x pass on a transformed column: the fields
// f = mean + fluctuation, V'(f) - V'(mean) in place of the spectrum (the k = 0
// mode of V'(mean) is added back in the z pass) and the sum of V in FP64.
template <class G, bool REC>
__device__ __forceinline__ void x_force_pointwise(cpx<typename G::T> (&e)[G::E], const typename G::T (&mean)[2], int nf,
                                                  const Couplings<typename G::T>& cp, double& vsum,
                                                  cpx<typename G::T>* __restrict__ rec_r, size_t base, size_t stride) {
    using T = typename G::T;
    const T inv = (T)(1.0 / (double)G::N3);
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const T d[2] = {e[m].x * inv, nf > 1 ? e[m].y * inv : (T)0};   // fluctuations
        const T f[2] = {mean[0] + d[0], mean[1] + d[1]};
        if constexpr (REC) rec_r[base + (size_t)G::TPL * m * stride] = cpx<T>{f[0], f[1]};  // adjoint recording
        // V'(f) - V'(mean) (dvdf, model.h:75-82), written so that every term
        // carries a fluctuation factor (no cancellation against the mean)
        T dv[2];
        dvdf_delta<Model, T>(mean, d, dv, cp);
        e[m].x = dv[0];
        e[m].y = nf > 1 ? dv[1] : (T)0;
        // V per site in the field precision, summed in FP64
        vsum += (double)potential_sum<Model, T>(nf, f, cp);
    }
}

// ---------------------------------------------------------------------------
// C: x-axis inverse, pointwise V'(phi, chi) of model.h (dvdf, model.h:75-82;
// kernel_field_k_nopad evolution.cpp:1103-1109) and x-axis forward, with the
// lattice sum of V (kernel_a_k_nopad evolution.cpp:1077-1085) in FP64.
// PADIN (the dealiased stepper on its padded grid): the x rows |x| > PH of the
// spectrum (PH = half the unpadded lattice) are zero and are neither read nor
// written.
template <class G, bool REC = false, bool PADIN = false, int PH = G::N / 4>
__global__ void __launch_bounds__((G::THREADS_COL)) __attribute__((amdgpu_waves_per_eu(1, G::XF_WAVES)))
k_x_force(cpx<typename G::T>* __restrict__ W, const MemberState* __restrict__ ms, const cpx<typename G::T>* __restrict__ tw,
          double* __restrict__ partV, const cpx<typename G::T>* __restrict__ Pf,
          cpx<typename G::T>* __restrict__ rec_r = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = G::PS;
    const size_t base = (size_t)member * G::MS + o * N + z + t * stride;
    cpx<T> e[G::E];
    auto kept = [&](int m) {   // spectrum row x = t + TPL m holds data
        const int x = t + G::TPL * m;
        return !PADIN || x <= PH || x >= N - PH;
    };
#pragma unroll
    for (int m = 0; m < G::E; ++m)
        e[m] = kept(m) ? W[base + (size_t)G::TPL * m * stride] : cpx<T>{(T)0, (T)0};
    // the member's couplings and the homogeneous values (the k = 0 mode left
    // out of the inverse transform) are read before the transform, so that
    // their latency is hidden behind it
    const Couplings<T> cp = couplings<T>(ms[member]);
    const int nf = ms[member].nflds;
    const cpx<T> dc = Pf[(size_t)member * G::MS];
    line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, tw);
    const T mean[2] = {(T)((double)dc.x / (double)G::N3), nf > 1 ? (T)((double)dc.y / (double)G::N3) : (T)0};
    double v[1] = {0.0};
    x_force_pointwise<G, REC>(e, mean, nf, cp, v[0], rec_r, base, stride);
    // In FP64 the forward transform reloads its twiddles (cached) instead of
    // keeping the inverse transform's live across the V' step: an opaque copy
    // of the table pointer.  That lowers the register count enough for two
    // blocks per CU at 512 points (measured 12 % faster); in FP32 the reuse is
    // the faster choice.
    const cpx<T>* tw_f = tw;
    if constexpr (sizeof(T) == 8) asm volatile("" : "+s"(tw_f));
    line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(e, t, c, lds, tw_f);
#pragma unroll
    for (int m = 0; m < G::E; ++m)
        if (kept(m)) W[base + (size_t)G::TPL * m * stride] = e[m];
    block_sum<1, G::THREADS_COL>(v, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0)
        partV[(size_t)member * G::BLOCKS_COL + blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

// C, multi-tile form for the long columns (512 points): each block loops over
// the tiles b, b + G, b + 2G, ... of its member (G = gridDim.x, one block per
// resident slot) and holds the next two tiles' columns in registers while it
// transforms the current one (the next one only where the columns are too
// long for two in the register budget, as at 768 points).  On gfx950 the block barriers of the transforms
// wait only on LDS, so the loads stay in flight across them: the load latency
// that one block per CU would otherwise expose is hidden behind the previous
// tiles' transforms.  The same arithmetic per column as k_x_force (bit-equal
// results); one V partial per block, at partV[member * gridDim.x + block].
template <class G, bool PADIN = false, int PH = G::N / 4>
__global__ void __launch_bounds__((G::THREADS_COL)) __attribute__((amdgpu_waves_per_eu(1, G::XF_WAVES)))
k_x_force_mt(cpx<typename G::T>* __restrict__ W, const MemberState* __restrict__ ms,
             const cpx<typename G::T>* __restrict__ tw, double* __restrict__ partV,
             const cpx<typename G::T>* __restrict__ Pf) {
    using T = typename G::T;
    constexpr int N = G::N;
    constexpr int NT = (N / G::TC) * N;   // tiles (TC columns x one y row) per member
    constexpr size_t stride = G::PS;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const int GS = (int)gridDim.x;
    auto tile_base = [&](int q) {
        return (size_t)member * G::MS + (size_t)(q / (N / G::TC)) * N + (size_t)(q % (N / G::TC)) * G::TC + c +
               (size_t)t * stride;
    };
    auto kept = [&](int m) {   // spectrum row x = t + TPL m holds data
        const int x = t + G::TPL * m;
        return !PADIN || x <= PH || x >= N - PH;
    };
    auto load = [&](cpx<T> (&dst)[G::E], int q) {
        const size_t b = tile_base(q);
#pragma unroll
        for (int m = 0; m < G::E; ++m) dst[m] = kept(m) ? W[b + (size_t)G::TPL * m * stride] : cpx<T>{(T)0, (T)0};
    };
    const Couplings<T> cp = couplings<T>(ms[member]);
    const int nf = ms[member].nflds;
    const cpx<T> dc = Pf[(size_t)member * G::MS];
    const T mean[2] = {(T)((double)dc.x / (double)G::N3), nf > 1 ? (T)((double)dc.y / (double)G::N3) : (T)0};
    double v = 0.0;
    // tiles prefetched: two while the three column sets (current and two
    // ahead) take at most 40 % of the VGPRs one block per CU may use
    // (measured: no spills for the 512- and 384-point shapes, spills at 768)
    constexpr int SET = G::E * (int)sizeof(cpx<T>) / 4;
    constexpr int BUDGET = G::THREADS_COL >= 256 ? 512 / (G::THREADS_COL / 256) : 512;
    constexpr int DEPTH = 3 * SET * 5 <= BUDGET * 2 ? 2 : 1;
    cpx<T> n1[G::E], n2[DEPTH > 1 ? G::E : 1];
    int q = blockIdx.x;
    if (q < NT) load(n1, q);
    if constexpr (DEPTH > 1)
        if (q + GS < NT) load(n2, q + GS);
    for (; q < NT; q += GS) {
        cpx<T> e[G::E];
        if constexpr (DEPTH > 1) {
#pragma unroll
            for (int m = 0; m < G::E; ++m) {
                e[m] = n1[m];
                n1[m] = n2[m];
            }
            if (q + 2 * GS < NT) load(n2, q + 2 * GS);
        } else {
#pragma unroll
            for (int m = 0; m < G::E; ++m) e[m] = n1[m];
            if (q + GS < NT) load(n1, q + GS);
        }
        line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, tw);
        x_force_pointwise<G, false>(e, mean, nf, cp, v, nullptr, 0, stride);
        line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(e, t, c, lds, tw);
        const size_t base = tile_base(q);
#pragma unroll
        for (int m = 0; m < G::E; ++m)
            if (kept(m)) W[base + (size_t)G::TPL * m * stride] = e[m];
    }
    double vv[1] = {v};
    block_sum<1, G::THREADS_COL>(vv, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) partV[(size_t)member * gridDim.x + blockIdx.x] = vv[0];
}

// The k = 0 mode of V' that the x pass leaves out: N^3 V'(pb, qb) with the
// homogeneous values pb + i qb = f(k = 0) / N^3, computed in FP64.
template <typename T>
__device__ __forceinline__ void add_mean_force(cpx<T>& vk, const cpx<T>& f0, const MemberState& s, double N3) {
    const Couplings<double> cp(s.par);
    const double mean[2] = {(double)f0.x / N3, s.nflds > 1 ? (double)f0.y / N3 : 0.0};
    vk.x += (T)(N3 * Model::dvdf<double>(0, mean, cp));
    if (s.nflds > 1) vk.y += (T)(N3 * Model::dvdf<double>(1, mean, cp));
}

// The homogeneous mode (k = 0) of the FP32 path, a compensated pair (the FP32
// array element plus the FP64 remainder in MemberState): the drift
// f += c_drift pi in FP64 (evolution.cpp:1164-1170).
template <typename T>
__device__ __forceinline__ void dc_drift(cpx<T>& f, const cpx<T>& p, MemberState& s) {
    const double fx = (double)f.x + s.dc_f[0] + s.c_drift * ((double)p.x + s.dc_p[0]);
    const double fy = (double)f.y + s.dc_f[1] + s.c_drift * ((double)p.y + s.dc_p[1]);
    f.x = (T)fx;
    f.y = (T)fy;
    s.dc_f[0] = fx - (double)f.x;
    s.dc_f[1] = fy - (double)f.y;
}

// ... and its two kicks in FP64: force F = -a^{3+alpha} (E + N^3 V'(mean)) with
// E the k = 0 residual of the transformed V' and the mean from the compensated
// field (n = 0: no Laplacian term); q1 = pi after the first kick, q2 after
// both (kernel_field_k_nopad evolution.cpp:1115-1127, kicks :1153-1158,
// :1184-1189).
template <typename T>
__device__ __forceinline__ void dc_kick(const cpx<T>& e0, const cpx<T>& f, const cpx<T>& p, const MemberState& s,
                                        double N3, double vc, double c1, double c2, double (&q1)[2],
                                        double (&q2)[2]) {
    const Couplings<double> cp(s.par);
    const double mean[2] = {((double)f.x + s.dc_f[0]) / N3, s.nflds > 1 ? ((double)f.y + s.dc_f[1]) / N3 : 0.0};
    const double F[2] = {-vc * ((double)e0.x + N3 * Model::dvdf<double>(0, mean, cp)),
                         -vc * ((double)e0.y + (s.nflds > 1 ? N3 * Model::dvdf<double>(1, mean, cp) : 0.0))};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        q1[i] = (i ? (double)p.y : (double)p.x) + s.dc_p[i] + c1 * F[i];
        q2[i] = q1[i] + c2 * F[i];
    }
}

// Both kicks of one field mode with the force F = -a^{1+alpha} kIR^2 n^2 f -
// a^{3+alpha} E (kernel_field_k_nopad evolution.cpp:1115-1127, kicks
// :1153-1158, :1184-1189); E is the mode's V' spectrum.  dc marks the
// homogeneous mode (its N^3 V'(mean) added in FP64; in the FP32 path its kicks
// on the compensated pair).  Between the kicks: the adjoint recording (REC),
// the synchronised store of a call's last kick (SYNC_OUT, streaming) and the
// sums of |pi|^2 and n^2 |f|^2 (REDUCE).  The caller stores p.
template <typename T, bool REC, bool SYNC_OUT, bool REDUCE>
__device__ __forceinline__ void kick_mode(cpx<T>& E, const cpx<T>& f, cpx<T>& p, int n2i, bool dc, MemberState& s,
                                          double N3, T lapc, T vc, double vcd, double c1d, double c2d,
                                          double (&red)[2], cpx<T>* rec_p, cpx<T>* ps) {
    const T c1 = (T)c1d, c2 = (T)c2d;
    if (sizeof(T) == 8 && dc) add_mean_force(E, f, s, N3);
    const bool dc32 = sizeof(T) == 4 && dc;
    double q1[2] = {0.0, 0.0}, q2[2] = {0.0, 0.0};
    if (dc32) dc_kick(E, f, p, s, N3, vcd, c1d, c2d, q1, q2);
    const T ln2 = lapc * (T)n2i;
    const cpx<T> F = {-ln2 * f.x - vc * E.x, -ln2 * f.y - vc * E.y};
    p.x += c1 * F.x;
    p.y += c1 * F.y;
    if (dc32) p = cpx<T>{(T)q1[0], (T)q1[1]};
    if constexpr (REC) *rec_p = p;   // adjoint recording: synchronised pi
    // synchronised pi at the end of a call, read only if the state is
    // synchronised later: a streaming store
    if constexpr (SYNC_OUT) {
        store_stream(ps, p);
        if (dc32) {
            s.dc_ps[0] = q1[0] - (double)p.x;
            s.dc_ps[1] = q1[1] - (double)p.y;
        }
    }
    if constexpr (REDUCE) {
        const double px_ = dc32 ? q1[0] : (double)p.x, py_ = dc32 ? q1[1] : (double)p.y;
        const double fx = f.x, fy = f.y;
        red[0] += px_ * px_ + py_ * py_;
        red[1] += (double)n2i * (fx * fx + fy * fy);
    }
    p.x += c2 * F.x;
    p.y += c2 * F.y;
    if (dc32) {
        p = cpx<T>{(T)q2[0], (T)q2[1]};
        s.dc_p[0] = q2[0] - (double)p.x;
        s.dc_p[1] = q2[1] - (double)p.y;
    }
}

// ---------------------------------------------------------------------------
// E: z-axis forward pass with the kick fused (kernel_field_k_nopad
// evolution.cpp:1115-1127 and the kicks :1153-1158, :1184-1189).  W holds the
// unnormalised spectrum of V'.  c1/c2 are w dt / 2 of the two kicks applied
// with this force; REDUCE accumulates sum |pi|^2 after the first kick and
// sum n^2 |f|^2 (kinetic_energy_k / gradient_energy_k, evolution.cpp:923-1035).
template <class G, bool REDUCE, bool REC = false, bool STREAM = false, bool SYNC_OUT = false>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_z_kick(const cpx<typename G::T>* __restrict__ W, const cpx<typename G::T>* __restrict__ Pf, cpx<typename G::T>* __restrict__ Pp,
         MemberState* __restrict__ ms, const cpx<typename G::T>* __restrict__ tw,
         double* __restrict__ partKG, KickArgs ka, StepConsts sc,
         cpx<typename G::T>* __restrict__ rec_p = nullptr, cpx<typename G::T>* __restrict__ Ps = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const int member = blockIdx.y;
    if (ka.prologue && ms[member].ahead) return;   // the call's first kick was applied as a look-ahead
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (size_t)member * G::MS + (line / N) * G::PS + (line % N) * N + t;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m)   // W is not read again; streamed in FP64 (measured 14 % faster
        // there, 1 % slower in FP32, where more of the arrays fit the memory-side cache)
        e[m] = (STREAM && sizeof(T) == 8) ? load_stream(W + base + G::TPL * m) : W[base + G::TPL * m];
    line_fft<T, N, typename G::Rad, -1, typename G::LayRow>(e, t, c, lds, tw);
    const double a = ka.prologue ? ms[member].a_start : ms[member].a;
    const T lapc = (T)(pow(a, 1 + sc.alpha) * sc.kIR * sc.kIR);
    const double vcd = pow(a, 3 + sc.alpha);
    const T vc = (T)vcd;
    double c1d, c2d;
    kick_coefs(ka, ms[member], c1d, c2d);
    const int ix = (int)(line / N), iy = (int)(line % N);
    const int px = signed_index(ix, N), py = signed_index(iy, N);
    const int pxy = px * px + py * py;
    double red[2] = {0.0, 0.0};
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const int kz = t + G::TPL * m;
        const int pz = signed_index(kz, N);
        const cpx<T> f = STREAM ? load_stream(Pf + base + G::TPL * m) : Pf[base + G::TPL * m];
        cpx<T> p = STREAM ? load_stream(Pp + base + G::TPL * m) : Pp[base + G::TPL * m];
        const size_t idx = base + G::TPL * m;
        kick_mode<T, REC, SYNC_OUT, REDUCE>(e[m], f, p, pxy + pz * pz, m == 0 && line == 0 && t == 0, ms[member],
                                            (double)G::N3, lapc, vc, vcd, c1d, c2d, red, REC ? rec_p + idx : nullptr,
                                            SYNC_OUT ? Ps + idx : nullptr);
        Pp[idx] = p;
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
    int prologue;      // the begin starts a call: members already ahead have begun
                       // it at the end of the last call and are left alone
    double w_finish, w_begin;
    double w_after;    // weight of the substep after the one begun (of the next
                       // step, w_0, when the one begun is the last)
    int begin_last;    // the substep begun is the last of its step
    int end_call;      // the finished substep ends a call: 1 with the look-ahead
                       // kick applied (ahead when its coefficient is nonzero), 2 without
    int nV, nKG;       // partial counts per member
};

// Merge of two compensated partial sums (Fast2Sum on the leading parts):
// symmetric in its arguments, so both lanes of an xor-shuffle pair agree.
__device__ __forceinline__ void comp_merge(double& s, double& c, double s2, double c2) {
    const double t = s + s2;
    const double err = fabs(s) >= fabs(s2) ? (s - t) + s2 : (s2 - t) + s;
    s = t;
    c = c + c2 + err;
}

// Threads per block of k_scalars: the partial sums of up to kScalarDirect
// blocks per pass are reduced directly, larger counts are first cut to
// chunk sums by k_prereduce.
constexpr int kScalarThreads = 256;
constexpr int kScalarDirect = 512;

static __global__ void __launch_bounds__(kScalarThreads)
k_scalars(MemberState* __restrict__ ms, const double* __restrict__ partV,
          const double* __restrict__ partKG, ScalarOp op, StepConsts sc, MemberState* __restrict__ rec = nullptr) {
    constexpr int WAVES = kScalarThreads / 64;
    __shared__ double sh[3 * 2 * WAVES];
    const int member = blockIdx.x;
    // the state is loaded up front (in flight during the reduction) and
    // written back once at the end
    MemberState s = ms[member];
    if (op.finish) {
        Comp acc[3];
        const double* pv = partV + (size_t)member * op.nV;
        const double* pk = partKG + (size_t)member * op.nKG * 2;
        // loads of a batch are all issued before the compensated adds, so a
        // batch costs one memory round trip
        constexpr int U = 4;
        const int n = op.nV > op.nKG ? op.nV : op.nKG;
        for (int i0 = threadIdx.x; i0 < n; i0 += U * kScalarThreads) {
            double v[U];
            double2 kg[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kScalarThreads;
                v[u] = i < op.nV ? pv[i] : 0.0;
                kg[u] = i < op.nKG ? reinterpret_cast<const double2*>(pk)[i] : double2{0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                acc[0].add(v[u]);
                acc[1].add(kg[u].x);
                acc[2].add(kg[u].y);
            }
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double s2 = __shfl_xor(acc[q].s, off, 64), c2 = __shfl_xor(acc[q].c, off, 64);
                comp_merge(acc[q].s, acc[q].c, s2, c2);
            }
            if (lane == 0) {
                sh[(q * WAVES + wave) * 2] = acc[q].s;
                sh[(q * WAVES + wave) * 2 + 1] = acc[q].c;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot[3];
            for (int q = 0; q < 3; ++q) {
                double ss = sh[(q * WAVES) * 2], cc = sh[(q * WAVES) * 2 + 1];
                for (int w = 1; w < WAVES; ++w) comp_merge(ss, cc, sh[(q * WAVES + w) * 2], sh[(q * WAVES + w) * 2 + 1]);
                tot[q] = ss + cc;
            }
            const double a = s.a;
            const double Kn = 0.5 * tot[1] / ipow(a, 6) / (sc.N3 * sc.N3);
            const double Gn = sc.kIR * sc.kIR * tot[2] / sc.N3 / 2.0 / (a * a) / sc.N3;
            const double Vn = tot[0] / sc.NV;
            s.K = Kn;
            s.G = Gn;
            s.V = Vn;
            const int al = sc.alpha;
            double add = 1.0 / 3.0 * sc.f2m * ipow(a, 1 + 2 * al) * ((al - 2) * Kn + al * Gn + (al + 1) * Vn);
            if (sc.expansion_static) add = 0.0;
            s.add = add;
            s.b = s.b_half + op.w_finish * s.dt / 2.0 * add;
            if (op.last_of_step) {
                s.ad = sc.expansion_static ? 0.0 : s.b;
                s.t += s.dt;
                if (s.dt != 0.0) s.steps += 1.0;
                s.dt = s.dt_next;
            }
            if (op.end_call) s.ahead = (op.end_call == 1 && s.kick2 != 0.0) ? 1 : 0;
        }
    }
    if (op.begin && threadIdx.x == 0 && !(op.prologue && s.ahead)) {
        if (op.prologue) s.dt = step_dt(s, s.a, s.t, sc);   // the first step of a call
        const double w = op.w_begin, dt = s.dt;
        s.b_half = s.b + w * dt / 2.0 * s.add;
        s.a_start = s.a;
        s.a_half = sc.expansion_static ? 1.0 : s.a + s.b_half * w * dt / 2.0;
        s.c_drift = w * dt * ipow(s.a_half, sc.alpha - 3);
        s.a = sc.expansion_static ? 1.0 : s.a_half + s.b_half * w * dt / 2.0;
        s.kick1 = w * dt / 2.0;
        if (op.begin_last) {   // the step ends with this substep: dt of the next step
            s.dt_next = step_dt(s, s.a, s.t + dt, sc);
            s.kick2 = op.w_after * s.dt_next / 2.0;
        } else {
            s.kick2 = op.w_after * dt / 2.0;
        }
    }
    if (threadIdx.x == 0) {
        ms[member] = s;
        if (rec) rec[member] = s;   // adjoint recording
    }
}

// ---------------------------------------------------------------------------
// Diagnostics: x-axis inverse pass that normalises by 1/N^3, writes the real
// fields (phi, chi) back into W, and reduces sum phi, sum chi, sum phi^2,
// sum chi^2 (meansvars output.cpp:427-447), the sum of each potential term
// (potential_energy, model.h:35-73) and of each field's effective mass
// (effective_mass, model.h:96-98, averaged as output.cpp:569-580): the
// kDiagWidth partial sums of a block.
constexpr int kDiagWidth = 4 + kMaxTerms + 2;
template <class G>
__global__ void __launch_bounds__((G::THREADS_COL))
k_x_inverse_diag(cpx<typename G::T>* __restrict__ W, const MemberState* __restrict__ ms,
                 const cpx<typename G::T>* __restrict__ tw, double* __restrict__ part,
                 const cpx<typename G::T>* __restrict__ src) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = G::PS;
    const size_t base = (size_t)member * G::MS + o * N + z + t * stride;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * stride];
    line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, tw);
    const double inv = 1.0 / (double)G::N3;
    const Couplings<double> cp(ms[member].par);
    const int nf = ms[member].nflds;
    const int nterms = Model::num_potential_terms(nf);
    const cpx<T> dc = src[(size_t)member * G::MS];   // k = 0 mode, left out of the z pass
    const double pb = (double)dc.x * inv, qb = (double)dc.y * inv;
    double r[kDiagWidth];
#pragma unroll
    for (int q = 0; q < kDiagWidth; ++q) r[q] = 0.0;
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const double phi = pb + (double)e[m].x * inv, chi = qb + (double)e[m].y * inv;
        e[m].x = (T)phi;
        e[m].y = (T)chi;
        r[0] += phi;
        r[1] += chi;
        r[2] += phi * phi;
        r[3] += chi * chi;
        const double f[2] = {phi, nf > 1 ? chi : 0.0};
#pragma unroll
        for (int term = 0; term < kMaxTerms; ++term)
            if (term < nterms) r[4 + term] += Model::potential_energy<double>(term, f, cp);
        double msq[2];
        Model::effective_mass<double>(msq, f, cp);
        r[4 + kMaxTerms] += msq[0];
        r[5 + kMaxTerms] += msq[1];
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + (size_t)G::TPL * m * stride] = e[m];
    block_sum<kDiagWidth, G::THREADS_COL>(r, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) {
        double* dst = part + ((size_t)member * G::BLOCKS_COL + blockIdx.y * gridDim.x + blockIdx.x) * kDiagWidth;
#pragma unroll
        for (int q = 0; q < kDiagWidth; ++q) dst[q] = r[q];
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
    constexpr size_t PS = plane_stride(N);
    const int member = blockIdx.y;
    const cpx<T>* f = Pf + (size_t)member * member_stride(N);
    const cpx<T>* p = Pp + (size_t)member * member_stride(N);
    double r[4] = {0, 0, 0, 0};
    for (size_t li = (size_t)blockIdx.x * 256 + threadIdx.x; li < N3; li += (size_t)blocks_per_member * 256) {
        const int i = (int)(li / ((size_t)N * N)), j = (int)((li / N) % N), k = (int)(li % N);
        const size_t idx = (size_t)i * PS + (size_t)j * N + k;
        const size_t midx = (size_t)((N - i) % N) * PS + (size_t)((N - j) % N) * N + (N - k) % N;
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
    constexpr int NH = N / 2 + 1;
    const int member = blockIdx.y;
    for (int i = threadIdx.x; i < nbins_total * 4; i += blockDim.x) hist[i] = 0.0;
    __syncthreads();
    constexpr size_t PS = plane_stride(N);
    const cpx<T>* f = Pf + (size_t)member * member_stride(N);
    const cpx<T>* p = Pp + (size_t)member * member_stride(N);
    const double a = ms[member].a, ad = ms[member].ad;
    const double aop = pow(a, alpha - 3), hub = ad / a;
    const size_t nk = (size_t)N * N * NH;
    for (size_t h = (size_t)blockIdx.x * blockDim.x + threadIdx.x; h < nk; h += (size_t)blocks_per_member * blockDim.x) {
        const int i = (int)(h / ((size_t)N * NH)), j = (int)((h / NH) % N), k = (int)(h % NH);
        const size_t idx = (size_t)i * PS + (size_t)j * N + k;
        const size_t midx = (size_t)((N - i) % N) * PS + (size_t)((N - j) % N) * N + (N - k) % N;
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

// First level of the substep reduction when a pass has many blocks: per
// member, the n partial vectors of width W are cut into `chunks` contiguous
// ranges, each summed with compensation in a fixed order.
template <int W>
__device__ void prereduce_chunk(const double* __restrict__ in, int n, double* __restrict__ out, int chunks,
                                int member, int chunk) {
    __shared__ double sh[W * 2 * 4];
    const long lo = (long)n * chunk / chunks, hi = (long)n * (chunk + 1) / chunks;
    Comp acc[W];
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x)
#pragma unroll
        for (int q = 0; q < W; ++q) acc[q].add(in[((size_t)member * n + i) * W + q]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < W; ++q) {
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
        for (int q = 0; q < W; ++q) {
            double ss = sh[(q * 4) * 2], cc = sh[(q * 4) * 2 + 1];
            for (int w = 1; w < 4; ++w) comp_merge(ss, cc, sh[(q * 4 + w) * 2], sh[(q * 4 + w) * 2 + 1]);
            out[((size_t)member * chunks + chunk) * W + q] = ss + cc;
        }
}

// Both first-level reductions of a substep in one launch: blocks
// [0, chunks) cut the V partials (width 1), blocks [chunks, 2 chunks) the
// K, G partials (width 2); a pass with few blocks (n <= chunks) is skipped.
static __global__ void __launch_bounds__(256)
k_prereduce(const double* __restrict__ inV, int nV, double* __restrict__ outV, const double* __restrict__ inKG,
            int nKG, double* __restrict__ outKG, int chunks) {
    const int member = blockIdx.y;
    if ((int)blockIdx.x < chunks) {
        if (nV > chunks) prereduce_chunk<1>(inV, nV, outV, chunks, member, blockIdx.x);
    } else if (nKG > chunks) {
        prereduce_chunk<2>(inKG, nKG, outKG, chunks, member, blockIdx.x - chunks);
    }
}

// Restores the synchronised state of the members a call left ahead by one
// look-ahead kick: the momenta from the sync buffer (k_sync_copy, `arrays`
// packed arrays per member, member stride `mstride`), then the scale factor
// a = a_start, the homogeneous mode's remainder of the synchronised pi and
// the flag (k_sync_state, after the copies).
template <typename T>
__global__ void __launch_bounds__(256)
k_sync_copy(cpx<T>* __restrict__ dst, const cpx<T>* __restrict__ src, const MemberState* __restrict__ ms,
            size_t n, size_t mstride, int arrays, int members) {
    const int member = blockIdx.y % members, arr = blockIdx.y / members;
    if (!ms[member].ahead) return;
    const size_t off = ((size_t)arr * members + member) * mstride;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[off + i] = src[off + i];
}
static __global__ void k_sync_state(MemberState* __restrict__ ms, int members) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < members && ms[m].ahead) {
        ms[m].a = ms[m].a_start;
        ms[m].dc_p[0] = ms[m].dc_ps[0];
        ms[m].dc_p[1] = ms[m].dc_ps[1];
        ms[m].ahead = 0;
    }
}

// Device-to-device copy of n complex elements, launched in place of a memcpy
// where the copy is part of a captured graph (so that the graph holds kernel
// nodes only).
template <typename T>
__global__ void __launch_bounds__(256)
k_copy(cpx<T>* __restrict__ dst, const cpx<T>* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// Fixed-order sum of per-block partial vectors: out[q] = sum_b part[b][q].
static __global__ void k_sum_partials(const double* __restrict__ part, int nblocks, int width,
                               double* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= width) return;
    Comp acc;
    for (int b = 0; b < nblocks; ++b) acc.add(part[(size_t)b * width + q]);
    out[q] = acc.value();
}

}  // namespace psc

// ===========================================================================
// Gravitational waves.  The 6 symmetric components u_ij (11,22,33,12,13,23) are
// packed in pairs U_0 = u11 + i u22, U_1 = u33 + i u12, U_2 = u13 + i u23 (and
// likewise pi_u), so the GW kick of the real-space stepper (kernel_GW,
// evolution.cpp:581-601, with lapl() evolution.cpp:343-372 as -k^2 u_k)
//   pi_u += c a^{1+alpha} (-kIR^2 n^2 U + 2 S_k)
// acts on packed spectra with real coefficients.  The source
// S_ij = sum_f d_i f d_j f uses the spectral first derivative with the Nyquist
// component zeroed (field_1st_derivative, evolution.cpp:466-578); the three
// packed gradients d_d phi + i d_d chi ride through the same passes as the
// fields in the work arrays W[1..3] (W[0] holds the fields / V').
namespace psc {

__device__ __forceinline__ int deriv_index(int i, int N) { return i == N / 2 ? 0 : signed_index(i, N); }

// Snapshot energy densities (snapshots, output.cpp:1614-1625), x-axis inverse
// pass into an x-major double array D (the order of the reference's
// f[fld][i][j][k]) holding the y rows o0 .. o0 + rows - 1 of every x plane:
// D[(x rows + r) N + z] is the site (x, o0 + r, z), so rows = N, o0 = 0 is the
// whole N^3 lattice.  MODE 0: potential V(phi, chi) summed over the terms
// (model.h:35-73; scale unused), with the k = 0 mode of src added back as in
// k_x_inverse_diag; MODE 1: scale * (pi_phi^2 + pi_chi^2), likewise;
// MODE 3: D = scale * (d phi^2 + d chi^2) for the first derivative axis and
// MODE 2: D += scale * (d phi^2 + d chi^2) for the others (no k = 0 mode).
// W and src are one member's arrays; the grid is (N / TC, rows).
template <class G, int MODE>
__global__ void __launch_bounds__((G::THREADS_COL))
k_x_density(const cpx<typename G::T>* __restrict__ W, double* __restrict__ D, const MemberState* __restrict__ ms,
            const cpx<typename G::T>* __restrict__ tw, const cpx<typename G::T>* __restrict__ src, double scale,
            int o0, int rows) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = (size_t)o0 + blockIdx.y;
    constexpr size_t stride = G::PS;
    const size_t base = o * N + z + t * stride;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * stride];
    line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, tw);
    const double inv = 1.0 / (double)G::N3;
    double pb = 0.0, qb = 0.0;
    if constexpr (MODE < 2) {
        const cpx<T> dc = src[0];
        pb = (double)dc.x * inv;
        qb = (double)dc.y * inv;
    }
    const Couplings<double> cp(ms[0].par);
    const int nf = ms[0].nflds;
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const double u = pb + (double)e[m].x * inv, v = nf > 1 ? qb + (double)e[m].y * inv : 0.0;
        const size_t x = (size_t)(t + G::TPL * m);
        double* d = D + (x * (size_t)rows + blockIdx.y) * N + z;
        if constexpr (MODE == 0) {
            const double f[2] = {u, v};
            *d = potential_sum<Model, double>(nf, f, cp);
        }
        else if constexpr (MODE == 1 || MODE == 3) *d = scale * (u * u + v * v);
        else *d += scale * (u * u + v * v);
    }
}

// Snapshot gradient density, first pass per axis: the spectral derivative
// i kIR n_d P_f along axis d (n_d = 0 at the Nyquist index, the 1D transforms
// of gradient_energy_lattice, evolution.cpp:23-138, restated in 3D) formed
// on the fly, then the z-axis inverse transform into W.  One member.
template <class G>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_z_inverse_deriv(const cpx<typename G::T>* __restrict__ Pf, cpx<typename G::T>* __restrict__ W,
                  const cpx<typename G::T>* __restrict__ tw, int d, double kIR) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (line / N) * G::PS + (line % N) * N + t;
    const int ix = (int)(line / N), iy = (int)(line % N);
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const int n = d == 0 ? deriv_index(ix, N) : (d == 1 ? deriv_index(iy, N) : deriv_index(t + G::TPL * m, N));
        const T k = (T)(kIR * n);
        const cpx<T> f = Pf[base + G::TPL * m];
        e[m] = cpx<T>{-k * f.y, k * f.x};
    }
    line_fft<T, N, typename G::Rad, +1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + G::TPL * m] = e[m];
}


// A (GW): drift of fields and (DRIFT_U) of u, then z-axis inverse of the
// fields (W0) and of the z gradient i kIR n_z P_f (W3).  STREAM: f and pi
// with non-temporal accesses (arrays beyond the memory-side cache; 1215 ->
// 1180 us at 512^3 FP32).  The x and y gradients need no z
// pass of their own: i kIR n_x and i kIR n_y are constant along a z line, so
// they are formed from W0 in the y and x passes (k_y_inverse_gw, k_x_gw).
template <class G, bool DRIFT, bool REC = false, bool DRIFT_U = true, bool STREAM = false>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_z_inverse_gw(cpx<typename G::T>* __restrict__ Pf, const cpx<typename G::T>* __restrict__ Pp, cpx<typename G::T>* __restrict__ W,
               cpx<typename G::T>* __restrict__ U, const cpx<typename G::T>* __restrict__ PU, MemberState* __restrict__ ms,
               const cpx<typename G::T>* __restrict__ tw, size_t astride, double kIR,
               cpx<typename G::T>* __restrict__ rec_f = nullptr, cpx<typename G::T>* __restrict__ rec_ph = nullptr,
               cpx<typename G::T>* __restrict__ rec_u = nullptr, cpx<typename G::T>* __restrict__ rec_puh = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const int member = blockIdx.y;
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (size_t)member * G::MS + (line / N) * G::PS + (line % N) * N + t;
    const T cd = DRIFT ? (T)ms[member].c_drift : (T)0;
    cpx<T> e[G::E];
    cpx<T> f[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        cpx<T> v = STREAM ? load_stream(Pf + base + G::TPL * m) : Pf[base + G::TPL * m];
        if constexpr (DRIFT) {
            const cpx<T> p = STREAM ? load_stream(Pp + base + G::TPL * m) : Pp[base + G::TPL * m];
            if constexpr (REC) {  // adjoint recording: substep start f and pi_half
                rec_f[base + G::TPL * m] = v;
                rec_ph[base + G::TPL * m] = p;
            }
            if (sizeof(T) == 4 && m == 0 && line == 0 && t == 0) {
                dc_drift(v, p, ms[member]);   // the homogeneous mode, compensated
            } else {
                v.x += cd * p.x;
                v.y += cd * p.y;
            }
            if constexpr (STREAM) store_stream(Pf + base + G::TPL * m, v);
            else Pf[base + G::TPL * m] = v;
        }
        f[m] = v;
        e[m] = v;
    }
    if (line == 0 && t == 0) e[0] = cpx<T>{(T)0, (T)0};   // k = 0 mode added back in k_x_force
    line_fft<T, N, typename G::Rad, +1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + G::TPL * m] = e[m];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const T kz = (T)(kIR * deriv_index(t + G::TPL * m, N));
        e[m] = cpx<T>{-kz * f[m].y, kz * f[m].x};
    }
    line_fft<T, N, typename G::Rad, +1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[3 * astride + base + G::TPL * m] = e[m];
    if constexpr (DRIFT && DRIFT_U) {   // else the u kick (k_z_kick_u) drifts u
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int m = 0; m < G::E; ++m) {
                const size_t idx = (size_t)a * astride + base + G::TPL * m;
                cpx<T> u = U[idx];
                const cpx<T> p = PU[idx];
                u.x += cd * p.x;
                u.y += cd * p.y;
                U[idx] = u;
                if constexpr (REC) {  // adjoint recording: u after the drift, pi_u half
                    rec_u[idx] = u;
                    rec_puh[idx] = p;
                }
            }
    }
}

// B (GW): y-axis inverse of the fields W0 and of the z gradient W3, and the
// y gradient i kIR n_y W0 formed from the fields' column before its
// transform, into W2.  W1 is not read (the x gradient comes from W0 in
// the x pass).  The fields' column is read twice (the second time from the
// cache) rather than held through the gradient's transform: one column in
// registers at a time keeps the 16-element plans free of spills.
template <class G>
__global__ void __launch_bounds__((G::THREADS_COL))
k_y_inverse_gw(cpx<typename G::T>* __restrict__ W, const cpx<typename G::T>* __restrict__ tw, size_t astride,
               double kIR) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t base = (size_t)member * G::MS + (size_t)blockIdx.y * G::PS + z + (size_t)t * N;
    cpx<T> e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const cpx<T> f = W[base + (size_t)G::TPL * m * N];
        const T ky = (T)(kIR * deriv_index(t + G::TPL * m, N));
        e[m] = cpx<T>{-ky * f.y, ky * f.x};
    }
    line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[2 * astride + base + (size_t)G::TPL * m * N] = e[m];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[base + (size_t)G::TPL * m * N];
    {   // twiddles reloaded (cached) rather than kept live across transforms
        const cpx<T>* twa = tw;
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, twa);
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + (size_t)G::TPL * m * N] = e[m];
#pragma unroll
    for (int m = 0; m < G::E; ++m) e[m] = W[3 * astride + base + (size_t)G::TPL * m * N];
    {   // twiddles reloaded (cached) rather than kept live across transforms
        const cpx<T>* twa = tw;
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, twa);
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[3 * astride + base + (size_t)G::TPL * m * N] = e[m];
}

// C (GW), first part: the x-axis inverse transforms of the three gradients
// (the x gradient i kIR n_x W0 formed from the fields' column, the y and z
// gradients from W2, W3), the packed source S = (S11 + i S22, S33 + i S12,
// S13 + i S23) from the real-space gradients, and the three x-axis forward
// transforms into W1..W3, in one pass.  It runs before k_x_force, which
// overwrites W0.
template <class G, bool REC = false>
__global__ void __launch_bounds__((G::THREADS_COL))
k_x_source(cpx<typename G::T>* __restrict__ W, const cpx<typename G::T>* __restrict__ tw, size_t astride,
           double kIR, cpx<typename G::T>* __restrict__ rec_g = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = G::PS;
    const size_t base = (size_t)member * G::MS + o * N + z + t * stride;
    const T inv = (T)(1.0 / (double)G::N3);
    cpx<T> g[3][G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const cpx<T> f = W[base + (size_t)G::TPL * m * stride];
        const T kx = (T)(kIR * deriv_index(t + G::TPL * m, N));
        g[0][m] = cpx<T>{-kx * f.y, kx * f.x};
    }
#pragma unroll
    for (int a = 1; a < 3; ++a)
#pragma unroll
        for (int m = 0; m < G::E; ++m) g[a][m] = W[(size_t)(a + 1) * astride + base + (size_t)G::TPL * m * stride];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const cpx<T>* twa = tw;   // reloaded per transform, as below
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(g[a], t, c, lds, twa);
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const T ax = g[0][m].x * inv, bx = g[0][m].y * inv;
        const T ay = g[1][m].x * inv, by = g[1][m].y * inv;
        const T az = g[2][m].x * inv, bz = g[2][m].y * inv;
        if constexpr (REC) {  // adjoint recording: real-space gradients d_d phi + i d_d chi
            const size_t idx = base + (size_t)G::TPL * m * stride;
            rec_g[idx] = cpx<T>{ax, bx};
            rec_g[astride + idx] = cpx<T>{ay, by};
            rec_g[2 * astride + idx] = cpx<T>{az, bz};
        }
        g[0][m] = cpx<T>{ax * ax + bx * bx, ay * ay + by * by};
        g[1][m] = cpx<T>{az * az + bz * bz, ax * ay + bx * by};
        g[2][m] = cpx<T>{ax * az + bx * bz, ay * az + by * bz};
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // each transform reloads its twiddles (cached) instead of keeping them
        // live across the three (an opaque copy of the table pointer): fewer
        // registers spilled at 512 points, 17 % faster there
        const cpx<T>* twa = tw;
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(g[a], t, c, lds, twa);
#pragma unroll
        for (int m = 0; m < G::E; ++m) W[(size_t)(a + 1) * astride + base + (size_t)G::TPL * m * stride] = g[a][m];
    }
}

// C (GW), both parts in one pass: per tile, the fields' column (W0) is read
// once; its x inverse, V' and x forward are those of k_x_force (W0 <- V'
// spectrum, the V partial sums at k_x_force's places), and the x gradient
// formed from the same column (kept in registers), with the y and z gradients
// (W2, W3), gives the source of k_x_source (W1..W3).  Moves 7X against 8X for
// the two kernels; the same arithmetic per column.
template <class G>
__global__ void __launch_bounds__((G::THREADS_COL))
k_x_gw(cpx<typename G::T>* __restrict__ W, const MemberState* __restrict__ ms, const cpx<typename G::T>* __restrict__ tw,
       double* __restrict__ partV, const cpx<typename G::T>* __restrict__ Pf, size_t astride, double kIR) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayCol::size(0)];
    const int c = threadIdx.x % G::TC, t = threadIdx.x / G::TC;
    const int member = blockIdx.z;
    const size_t z = (size_t)blockIdx.x * G::TC + c;
    const size_t o = blockIdx.y;
    constexpr size_t stride = G::PS;
    const size_t base = (size_t)member * G::MS + o * N + z + t * stride;
    cpx<T> f0[G::E], e[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        f0[m] = W[base + (size_t)G::TPL * m * stride];
        e[m] = f0[m];
    }
    const Couplings<T> cp = couplings<T>(ms[member]);
    const int nf = ms[member].nflds;
    const cpx<T> dc = Pf[(size_t)member * G::MS];
    // the fields: x inverse, V' - V'(mean), x forward (k_x_force)
    {
        const cpx<T>* twa = tw;
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(e, t, c, lds, twa);
    }
    const T mean[2] = {(T)((double)dc.x / (double)G::N3), nf > 1 ? (T)((double)dc.y / (double)G::N3) : (T)0};
    double v[1] = {0.0};
    x_force_pointwise<G, false>(e, mean, nf, cp, v[0], nullptr, 0, stride);
    {
        const cpx<T>* twa = tw;
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(e, t, c, lds, twa);
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) W[base + (size_t)G::TPL * m * stride] = e[m];
    // the source from the x gradient of the same column and the y, z gradients
    // (k_x_source)
    cpx<T> g[3][G::E];
#pragma unroll
    for (int a = 1; a < 3; ++a)
#pragma unroll
        for (int m = 0; m < G::E; ++m) g[a][m] = W[(size_t)(a + 1) * astride + base + (size_t)G::TPL * m * stride];
    const T inv = (T)(1.0 / (double)G::N3);
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const T kx = (T)(kIR * deriv_index(t + G::TPL * m, N));
        g[0][m] = cpx<T>{-kx * f0[m].y, kx * f0[m].x};
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const cpx<T>* twa = tw;
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, +1, typename G::LayCol>(g[a], t, c, lds, twa);
    }
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const T ax = g[0][m].x * inv, bx = g[0][m].y * inv;
        const T ay = g[1][m].x * inv, by = g[1][m].y * inv;
        const T az = g[2][m].x * inv, bz = g[2][m].y * inv;
        g[0][m] = cpx<T>{ax * ax + bx * bx, ay * ay + by * by};
        g[1][m] = cpx<T>{az * az + bz * bz, ax * ay + bx * by};
        g[2][m] = cpx<T>{ax * az + bx * bz, ay * az + by * bz};
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const cpx<T>* twa = tw;
        asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, -1, typename G::LayCol>(g[a], t, c, lds, twa);
#pragma unroll
        for (int m = 0; m < G::E; ++m) W[(size_t)(a + 1) * astride + base + (size_t)G::TPL * m * stride] = g[a][m];
    }
    block_sum<1, G::THREADS_COL>(v, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0)
        partV[(size_t)member * G::BLOCKS_COL + blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

// E (GW): z-axis forward of V' and of the source, field kick as k_z_kick and
// the GW kick of the three packed u arrays (DRIFT_U: after the substep's u
// drift, as k_z_kick_u; the recording replay of the adjoint also records u
// after the drift and pi_u half here).
template <class G, bool REDUCE, bool REC = false, bool SYNC_OUT = false, bool DRIFT_U = false>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_z_kick_gw(const cpx<typename G::T>* __restrict__ W, const cpx<typename G::T>* __restrict__ Pf, cpx<typename G::T>* __restrict__ Pp,
            cpx<typename G::T>* __restrict__ U, cpx<typename G::T>* __restrict__ PU, MemberState* __restrict__ ms,
            const cpx<typename G::T>* __restrict__ tw, double* __restrict__ partKG, KickArgs ka,
            StepConsts sc, size_t astride, cpx<typename G::T>* __restrict__ rec_p = nullptr,
            cpx<typename G::T>* __restrict__ Ps = nullptr, cpx<typename G::T>* __restrict__ PUs = nullptr,
            cpx<typename G::T>* __restrict__ rec_u = nullptr, cpx<typename G::T>* __restrict__ rec_puh = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const int member = blockIdx.y;
    if (ka.prologue && ms[member].ahead) return;   // the call's first kick was applied as a look-ahead
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (size_t)member * G::MS + (line / N) * G::PS + (line % N) * N + t;
    const double a = ka.prologue ? ms[member].a_start : ms[member].a;
    const T norm1 = (T)pow(a, 1 + sc.alpha);
    const T lapc = (T)(pow(a, 1 + sc.alpha) * sc.kIR * sc.kIR);
    const double vcd = pow(a, 3 + sc.alpha);
    const T vc = (T)vcd;
    const T k2c = (T)(sc.kIR * sc.kIR);
    double c1d, c2d;
    kick_coefs(ka, ms[member], c1d, c2d);
    const T c1 = (T)c1d, c2 = (T)c2d;
    const T cd = DRIFT_U ? (T)ms[member].c_drift : (T)0;
    const int ix = (int)(line / N), iy = (int)(line % N);
    const int px = signed_index(ix, N), py = signed_index(iy, N);
    const int pxy = px * px + py * py;
    double red[2] = {0.0, 0.0};
    // The field and momentum (and below u, pi_u) loads are issued before the
    // transform they follow, so their latency overlaps its LDS exchanges
    // (measured 10 % faster at 512^3, 7 % at 256^3; the same change in the
    // single-array k_z_kick measured slower).
    cpx<T> e[G::E], fa[G::E], pa[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        e[m] = W[base + G::TPL * m];
        fa[m] = Pf[base + G::TPL * m];
        pa[m] = Pp[base + G::TPL * m];
    }
    line_fft<T, N, typename G::Rad, -1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const int pz = signed_index(t + G::TPL * m, N);
        const cpx<T> f = fa[m];
        cpx<T> p = pa[m];
        const size_t idx = base + G::TPL * m;
        kick_mode<T, REC, SYNC_OUT, REDUCE>(e[m], f, p, pxy + pz * pz, m == 0 && line == 0 && t == 0, ms[member],
                                            (double)G::N3, lapc, vc, vcd, c1d, c2d, red, REC ? rec_p + idx : nullptr,
                                            SYNC_OUT ? Ps + idx : nullptr);
        Pp[idx] = p;
    }
#pragma unroll
    for (int ga = 0; ga < 3; ++ga) {
#pragma unroll
        for (int m = 0; m < G::E; ++m) {
            const size_t idx = (size_t)ga * astride + base + G::TPL * m;
            e[m] = W[(size_t)(ga + 1) * astride + base + G::TPL * m];
            fa[m] = U[idx];
            pa[m] = PU[idx];
        }
        // with 16 elements per thread the twiddles are reloaded per transform
        // (see k_x_source): measured 11 % faster at 256 points, 4 % slower
        // with the 8-element 512-point plan, which keeps them live
        const cpx<T>* twa = tw;
        if constexpr (G::E >= 16) asm volatile("" : "+s"(twa));
        line_fft<T, N, typename G::Rad, -1, typename G::LayRow>(e, t, c, lds, twa);
#pragma unroll
        for (int m = 0; m < G::E; ++m) {
            const int pz = signed_index(t + G::TPL * m, N);
            const T k2 = k2c * (T)(pxy + pz * pz);
            const size_t idx = (size_t)ga * astride + base + G::TPL * m;
            cpx<T> u = fa[m];
            cpx<T> p = pa[m];
            if constexpr (DRIFT_U) {   // the substep's u drift, as k_z_kick_u
                u.x += cd * p.x;
                u.y += cd * p.y;
                U[idx] = u;
                if constexpr (REC) {   // adjoint recording: u after the drift, pi_u half
                    rec_u[idx] = u;
                    rec_puh[idx] = p;
                }
            }
            const cpx<T> F = {norm1 * (-k2 * u.x + (T)2 * e[m].x), norm1 * (-k2 * u.y + (T)2 * e[m].y)};
            p.x += c1 * F.x;
            p.y += c1 * F.y;
            if constexpr (SYNC_OUT) store_stream(PUs + idx, p);
            p.x += c2 * F.x;
            p.y += c2 * F.y;
            PU[idx] = p;
        }
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

// E (GW), the u part alone: z-axis forward of one packed source array
// (W1..W3) and the GW kick of its u pair, one array per block (blockIdx.z);
// the fields' kick runs beside it as k_z_kick.  Same arithmetic as the u part
// of k_z_kick_gw, at lower register pressure (measured 3301 -> 3171 us for
// the whole kick at 512^3 FP32; k_z_kick_gw stays for the adjoint's
// recording replay).  DRIFT applies the substep's drift of u first, in place
// of k_z_inverse_gw: nothing reads u between the two, and the kick loads u
// and pi_u anyway, so the drift costs one store instead of two loads and a
// store in the z inverse (50X -> 44X per substep).  STREAM reads and writes the
// arrays with non-temporal accesses (arrays beyond the memory-side cache:
// 2986 -> 2782 us at 512^3 FP32).
template <class G, bool SYNC_OUT = false, bool DRIFT = false, bool STREAM = false>
__global__ void __launch_bounds__((G::THREADS_ROW))
k_z_kick_u(const cpx<typename G::T>* __restrict__ W, cpx<typename G::T>* __restrict__ U,
           cpx<typename G::T>* __restrict__ PU, const MemberState* __restrict__ ms,
           const cpx<typename G::T>* __restrict__ tw, KickArgs ka, StepConsts sc, size_t astride,
           cpx<typename G::T>* __restrict__ PUs = nullptr) {
    using T = typename G::T;
    constexpr int N = G::N;
    __shared__ cpx<T> lds[G::LayRow::size(G::L)];
    const int t = threadIdx.x % G::TPL, c = threadIdx.x / G::TPL;
    const int member = blockIdx.y;
    const int ga = blockIdx.z;
    if (ka.prologue && ms[member].ahead) return;   // the call's first kick was applied as a look-ahead
    const size_t line = (size_t)blockIdx.x * G::L + c;
    const size_t base = (size_t)member * G::MS + (line / N) * G::PS + (line % N) * N + t;
    const double a = ka.prologue ? ms[member].a_start : ms[member].a;
    const T norm1 = (T)pow(a, 1 + sc.alpha);
    const T k2c = (T)(sc.kIR * sc.kIR);
    double c1d, c2d;
    kick_coefs(ka, ms[member], c1d, c2d);
    const T c1 = (T)c1d, c2 = (T)c2d;
    const int px = signed_index((int)(line / N), N), py = signed_index((int)(line % N), N);
    const int pxy = px * px + py * py;
    const T cd = DRIFT ? (T)ms[member].c_drift : (T)0;
    cpx<T> e[G::E], fa[G::E], pa[G::E];
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const size_t idx = (size_t)ga * astride + base + G::TPL * m;
        e[m] = STREAM ? load_stream(W + (size_t)(ga + 1) * astride + base + G::TPL * m)
                      : W[(size_t)(ga + 1) * astride + base + G::TPL * m];
        fa[m] = STREAM ? load_stream(U + idx) : U[idx];
        pa[m] = STREAM ? load_stream(PU + idx) : PU[idx];
    }
    line_fft<T, N, typename G::Rad, -1, typename G::LayRow>(e, t, c, lds, tw);
#pragma unroll
    for (int m = 0; m < G::E; ++m) {
        const int pz = signed_index(t + G::TPL * m, N);
        const T k2 = k2c * (T)(pxy + pz * pz);
        const size_t idx = (size_t)ga * astride + base + G::TPL * m;
        cpx<T> u = fa[m];
        cpx<T> p = pa[m];
        if constexpr (DRIFT) {   // the substep's drift of u (evolution.cpp:654-663)
            u.x += cd * p.x;
            u.y += cd * p.y;
            if constexpr (STREAM) store_stream(U + idx, u);
            else U[idx] = u;
        }
        const cpx<T> F = {norm1 * (-k2 * u.x + (T)2 * e[m].x), norm1 * (-k2 * u.y + (T)2 * e[m].y)};
        p.x += c1 * F.x;
        p.y += c1 * F.y;
        if constexpr (SYNC_OUT) store_stream(PUs + idx, p);
        p.x += c2 * F.x;
        p.y += c2 * F.y;
        if constexpr (STREAM) store_stream(PU + idx, p);
        else PU[idx] = p;
    }
}

// GW spectrum (output.cpp:780-1062): modes visited as the reference does (the
// (+x) half of the kz = 0, N/2 planes and all interior kz with weight 2, real
// corners with weight 1, no DC), components unpacked from the pairs, one-sided
// P.u (projection 0) or Lambda u (projection 1).  Histograms: linear bins
// l = floor(|n| + 0.5) of (numpoints, h2, hdot2) and nlog logarithmic bins of
// hdot2 on ln|n| in [0, ln(sqrt(3) N / 2)].
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_gw_spectra(const cpx<T>* __restrict__ U, const cpx<T>* __restrict__ PU, size_t astride, int projection,
             int nbins_total, int nlog, double* __restrict__ part, int blocks) {
    extern __shared__ double hist[];  // 3 * nbins_total + nlog
    constexpr int NH = N / 2 + 1;
    const int width = 3 * nbins_total + nlog;
    for (int i = threadIdx.x; i < width; i += blockDim.x) hist[i] = 0.0;
    __syncthreads();
    const size_t nk = (size_t)N * N * NH;
    constexpr size_t PS = plane_stride(N);
    const double lmax = log(sqrt(3.0) * (N / 2));
    for (size_t h = (size_t)blockIdx.x * blockDim.x + threadIdx.x; h < nk; h += (size_t)blocks * blockDim.x) {
        const int i = (int)(h / ((size_t)N * NH)), j = (int)((h / NH) % N), k = (int)(h % NH);
        double wgt;
        if (k != 0 && k != N / 2) {
            wgt = 2.0;
        } else {
            const bool iedge = (i == 0 || i == N / 2), jedge = (j == 0 || j == N / 2);
            if (iedge && jedge) {
                if (i == 0 && j == 0 && k == 0) continue;
                wgt = 1.0;
            } else if (iedge) {
                if (!(j >= 1 && j < N / 2)) continue;
                wgt = 2.0;
            } else if (i >= 1 && i < N / 2) {
                wgt = 2.0;
            } else {
                continue;
            }
        }
        const size_t idx = (size_t)i * PS + (size_t)j * N + k;
        const size_t midx = (size_t)((N - i) % N) * PS + (size_t)((N - j) % N) * N + (N - k) % N;
        double ur[6], ui[6], pr[6], pi_[6];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const cpx<T> A = U[a * astride + idx], B = U[a * astride + midx];
            const double ar = A.x, ai = A.y, br = B.x, bi = -(double)B.y;
            ur[2 * a] = 0.5 * (ar + br); ui[2 * a] = 0.5 * (ai + bi);
            ur[2 * a + 1] = 0.5 * (ai - bi); ui[2 * a + 1] = -0.5 * (ar - br);
            const cpx<T> C = PU[a * astride + idx], D = PU[a * astride + midx];
            const double cr = C.x, ci = C.y, dr = D.x, di = -(double)D.y;
            pr[2 * a] = 0.5 * (cr + dr); pi_[2 * a] = 0.5 * (ci + di);
            pr[2 * a + 1] = 0.5 * (ci - di); pi_[2 * a + 1] = -0.5 * (cr - dr);
        }
        const int px = signed_index(i, N), py = signed_index(j, N), pz = k;
        const double kn = sqrt((double)(px * px + py * py + pz * pz));
        const double kh[3] = {px / kn, py / kn, pz / kn};
        double P[3][3];
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y) P[x][y] = (x == y ? 1.0 : 0.0) - kh[x] * kh[y];
        // component index of (x, y) in 11,22,33,12,13,23 order
        const int cix[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
        double h2v[2];
#pragma unroll
        for (int which = 0; which < 2; ++which) {
            const double* mr = which == 0 ? ur : pr;
            const double* mi = which == 0 ? ui : pi_;
            double vr[3][3], vi[3][3];
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) {
                    double sr = 0, si = 0;
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        sr += P[x][q] * mr[cix[q][y]];
                        si += P[x][q] * mi[cix[q][y]];
                    }
                    vr[x][y] = sr;
                    vi[x][y] = si;
                }
            double s = 0.0;
            if (projection == 0) {
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int y = 0; y < 3; ++y) s += vr[x][y] * vr[x][y] + vi[x][y] * vi[x][y];
                const double tr = vr[0][0] + vr[1][1] + vr[2][2], ti = vi[0][0] + vi[1][1] + vi[2][2];
                s -= (tr * tr + ti * ti) / 2.0;
            } else {
                const double tr = vr[0][0] + vr[1][1] + vr[2][2], ti = vi[0][0] + vi[1][1] + vi[2][2];
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int y = 0; y < 3; ++y) {
                        double ar = 0, ai = 0;
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            ar += vr[x][q] * P[q][y];
                            ai += vi[x][q] * P[q][y];
                        }
                        ar -= 0.5 * P[x][y] * tr;
                        ai -= 0.5 * P[x][y] * ti;
                        s += ar * ar + ai * ai;
                    }
            }
            h2v[which] = s;
        }
        const int l = (int)floor(kn + 0.5);
        atomicAdd(&hist[l * 3 + 0], wgt);
        atomicAdd(&hist[l * 3 + 1], wgt * h2v[0]);
        atomicAdd(&hist[l * 3 + 2], wgt * h2v[1]);
        int lb = (int)floor(log(kn) / lmax * nlog);
        lb = lb < 0 ? 0 : (lb >= nlog ? nlog - 1 : lb);
        atomicAdd(&hist[3 * nbins_total + lb], wgt * h2v[1]);
    }
    __syncthreads();
    double* dst = part + (size_t)blockIdx.x * width;
    for (int i = threadIdx.x; i < width; i += blockDim.x) dst[i] = hist[i];
}

}  // namespace psc
