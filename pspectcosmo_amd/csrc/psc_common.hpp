// psc_common.hpp -- shared device/host definitions of the lattice step.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include "psc_fft.hpp"

namespace psc {

// Per-member background state in device memory (FP64).  t, a, ad, add mirror
// the reference globals (pspectcosmo.h:44-45); b, b_half, a_start, a_half and
// c_drift are the temporaries of evolve_VV_k_nopad (evolution.cpp:1133-1194)
// carried across kernel launches so a captured step needs no host round trip.
struct MemberState {
    double t, a, ad, add;
    double b, b_half, a_start, a_half;
    double c_drift;
    double K, G, V;   // reductions of the last completed substep (normalised)
    double m2, g2;    // model couplings (model.h)
    double pad0, pad1;
};
static_assert(sizeof(MemberState) == 128, "MemberState layout");

// Constants of a run shared by all members.
struct StepConsts {
    double dt;
    double kIR;
    double f2m;       // (fStar / MPl)^2
    double N3;        // N^3
    int alpha;
    int expansion_static;
    int nflds;
    int pad;
};

// Geometry of the FFT passes for (T, N).
template <typename T, int N>
struct Geo {
    static constexpr int TPL = Plan<N>::TPL;
    static constexpr int E = N / TPL;            // elements per thread
    static constexpr int R1 = N / TPL;           // first radix
    // columns per block of the strided passes (lanes run along z)
    static constexpr int TC = sizeof(T) == 4 ? (N >= 512 ? 16 : (N >= 32 ? 32 : N))
                                             : (N >= 512 ? 8 : (N >= 128 ? 16 : (N >= 32 ? 32 : N)));
    // lines per block of the contiguous (z) passes
    static constexpr int L = sizeof(T) == 4 ? (N >= 512 ? 4 : (N >= 256 ? 16 : (N >= 64 ? 32 : 64)))
                                            : (N >= 512 ? 4 : (N >= 256 ? 8 : (N >= 128 ? 16 : (N >= 64 ? 32 : 64))));
    static constexpr int THREADS_COL = TC * TPL;
    static constexpr int THREADS_ROW = L * TPL;
    static constexpr size_t N3 = (size_t)N * N * N;
    static constexpr int BLOCKS_COL = (N / TC) * N;   // per member
    static constexpr int BLOCKS_ROW = (N * N) / L;    // per member
    using LayRow = LayoutRow<N, R1>;
    using LayCol = LayoutCol<N, TC>;
};

__device__ __forceinline__ int signed_index(int i, int N) { return i <= N / 2 ? i : i - N; }

// Block-wide sum of NV doubles, deterministic (fixed shuffle and LDS order).
// Result valid in thread 0.
template <int NV, int THREADS>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch /* >= NV*THREADS/64 */) {
    constexpr int WAVES = (THREADS + 63) / 64;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) scratch[q * WAVES + wave] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < WAVES; ++w) s += scratch[q * WAVES + w];
            v[q] = s;
        }
    }
}

// Neumaier-compensated accumulator.
struct Comp {
    double s = 0.0, c = 0.0;
    __host__ __device__ void add(double x) {
        double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __host__ __device__ void merge(const Comp& o) {
        add(o.s);
        c += o.c;
    }
    __host__ __device__ double value() const { return s + c; }
};

}  // namespace psc
