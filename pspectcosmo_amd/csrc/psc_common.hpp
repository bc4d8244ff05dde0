// psc_common.hpp -- shared device/host definitions of the lattice step.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include "../../include/psc.h"
#include "psc_fft.hpp"
#include "psc_model.hpp"

namespace psc {

// Per-member background state in device memory (FP64).  t, a, ad, add mirror
// the reference globals (pspectcosmo.h:44-45); b, b_half, a_start, a_half and
// c_drift are the temporaries of evolve_VV_k_nopad (evolution.cpp:1133-1194)
// carried across kernel launches so a captured step needs no host round trip.
// The step size and the kick coefficients live here too, so that a captured
// step does not depend on dt and the reference's per-step dt rule
// (calculate_dt, pspectcosmo.cpp:213-233, 622-627) runs on the device.
struct MemberState {
    double t, a, ad, add;
    double b, b_half, a_start, a_half;
    double c_drift;
    double K, G, V;   // reductions of the last completed substep (normalised)
    int nflds;        // fields evolved (1 or 2; with one, chi is held at 0)
    int ahead;        // pi (and pi_u, the ghosts) hold one look-ahead kick: the
                      // synchronised values are in the sync buffers, the
                      // synchronised a is a_start
    double dt;        // the step being taken
    double kick1;     // the coming z kick's coefficients: its substep's second
    double kick2;     // kick w dt / 2, and the next substep's first (the next
                      // step's when the substep is the last: the look-ahead)
    psc_model par;    // couplings of the model (model.h)
    double dt_next;   // dt of the next step (set when the last substep begins)
    double t_out;     // adaptive rule: the output time the steps land on
    double dt_cap;    // adaptive rule: largest dt (tOutputFreq)
    double steps;     // steps taken with dt > 0 since the data were loaded
    // FP32 path: the homogeneous mode (k = 0) of f and pi is a compensated
    // pair, the FP32 array element plus these FP64 remainders (re, im), so
    // that thousands of drifts and kicks do not accumulate FP32 rounding in
    // it; dc_ps belongs to the synchronised pi of a state left ahead.  Zero
    // in the FP64 path.
    double dc_f[2], dc_p[2], dc_ps[2];
    double reserved[6];
};
static_assert(sizeof(MemberState) == 320, "MemberState layout");

// Kick coefficients of a z-kick launch.
struct KickArgs {
    double c1, c2;      // explicit coefficients (from_state == 0: the adjoint's replays)
    int from_state;     // 1: c1, c2 = the member's kick1, kick2
    int prologue;       // first kick of a call's first substep only, at a_start
                        // (c2 = 0); members already ahead are left alone
    int no_lookahead;   // c2 = 0: the last substep of a call that ends synchronised
    int pad;
};

// The coefficients a kick launch applies to one member.
__device__ __forceinline__ void kick_coefs(const KickArgs& ka, const MemberState& s, double& c1, double& c2) {
    c1 = ka.from_state ? s.kick1 : ka.c1;
    c2 = (ka.prologue || ka.no_lookahead) ? 0.0 : (ka.from_state ? s.kick2 : ka.c2);
}

// The couplings of a member in the field precision.
template <typename T>
__device__ __forceinline__ Couplings<T> couplings(const MemberState& s) {
    return Couplings<T>(s.par);
}

// Constants of a run shared by all members.
struct StepConsts {
    double dt;        // fixed step (PSC_DT_FIXED)
    double kIR;
    double f2m;       // (fStar / MPl)^2
    double N3;        // N^3
    int alpha;
    int expansion_static;
    int nflds;
    int dt_rule;      // PSC_DT_FIXED or PSC_DT_REFERENCE
    double NV;        // sites of the lattice V is averaged over (N^3, or N_pad^3 dealiased)
    double kmax;      // kIR sqrt(3) N / 2, the momentum of calculate_dt (pspectcosmo.cpp:622)
};

// dt of a step that starts at (a, t): the fixed step, or the reference's rule
// dt = 2 / (kmax / a) / 10, capped at tOutputFreq and clamped to land on the
// next output time (pspectcosmo.cpp:213-233, 622-627); a member that has
// reached its output time (within the reference's 1e-10) takes steps of 0.
__host__ __device__ inline double step_dt(const MemberState& s, double a, double t, const StepConsts& sc) {
    if (sc.dt_rule == 0) return s.dt;
    double dt = 2.0 / (sc.kmax / a) / 10.0;
    dt = dt < s.dt_cap ? dt : s.dt_cap;
    if (t + dt > s.t_out) {
        const double rem = s.t_out - t;
        dt = rem < 1e-10 ? 0.0 : rem;
    }
    return dt;
}

// Launch shape of the FFT passes: precision T, lattice N, radix plan Rad
// (first radix R1, TPL = N / R1 threads per line), TC consecutive z columns
// per block in the strided passes, L lines per block in the contiguous passes.
template <typename T_, int N_, class Rad_, int TC_, int L_>
struct Shape {
    using T = T_;
    using Rad = Rad_;
    static constexpr int N = N_;
    static constexpr int R1 = first_radix(Rad{});
    static constexpr int TPL = N / R1;
    static constexpr int E = R1;                  // elements per thread
    static constexpr int TC = TC_;
    static constexpr int L = L_;
    static constexpr int THREADS_COL = TC * TPL;
    static constexpr int THREADS_ROW = L * TPL;
    static constexpr size_t N3 = (size_t)N * N * N;
    // Device arrays pad every x-plane with one unused row: plane stride
    // N (N + 1) elements.  A power-of-two plane stride (2 MiB at 512^3 FP32)
    // concentrates the x-axis passes' accesses on few DRAM channels.
    static constexpr size_t PS = (size_t)N * (N + 1);
    static constexpr size_t MS = (size_t)N * PS;      // member stride
    static constexpr int BLOCKS_COL = (N / TC) * N;   // per member
    // Most waves per SIMD k_x_force may run at.  At N <= 256 in FP32 the cap
    // of 3 lets the compiler keep more of the two transforms in registers
    // (130 VGPRs instead of 114): measured 11-12 % faster x passes for the
    // 16-column shapes and 7 % for r16.8_tc32 at 128, no change for the
    // others; FP64 and N = 512 are unchanged by any cap (and at 512 a cap
    // would cost the second block per CU), so they keep none (8).
    static constexpr int XF_WAVES = (sizeof(T_) == 4 && N_ <= 256) ? 3 : 8;
    // The force x pass runs in its multi-tile form (k_x_force_mt: blocks loop
    // over tiles, the next two tiles' columns prefetched into registers) where
    // one block fills a CU: measured 23 % faster at 512 points FP32 (and 15-30 %
    // for the padded 512- and 384-point passes and the plain 384 one), 6 % in
    // FP64 at 512; no gain at 256 points and below, where several blocks per CU
    // already overlap their loads with each other's transforms.
    static constexpr bool XMT = sizeof(T_) == 4 ? N_ >= 384 : N_ >= 512;
    static constexpr int BLOCKS_ROW = (N * N) / L;    // per member
    using LayRow = LayoutRow<N, R1>;
    using LayCol = LayoutCol<N, TC>;
    static_assert(THREADS_COL <= 1024 && THREADS_ROW <= 1024, "block too large");
    static_assert(TC >= 8 && L >= 4, "the partial-sum buffers hold (N / 8) N column and N^2 / 4 row blocks");
    static_assert(LayCol::size(0) * sizeof(cpx<T>) <= 131072, "strided tile exceeds 128 KiB of LDS");
    static_assert(LayRow::size(L) * sizeof(cpx<T>) <= 65536, "row tile exceeds 64 KiB of LDS");
};

__device__ __forceinline__ int signed_index(int i, int N) { return i <= N / 2 ? i : i - N; }

// Non-temporal (streaming) access for data the next passes do not read:
// such lines then leave room in the caches for the arrays the next pass
// does read.
template <typename T>
__device__ __forceinline__ cpx<T> load_stream(const cpx<T>* p) {
    using V = T __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    return cpx<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void store_stream(cpx<T>* p, cpx<T> a) {
    using V = T __attribute__((ext_vector_type(2)));
    V v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
}

// a^n for a small integer exponent (powers of the scale factor) by repeated
// squaring: a few multiplications instead of a general pow on the single
// thread that runs each scalar update.
__device__ __forceinline__ double ipow(double a, int n) {
    double r = 1.0, b = n < 0 ? 1.0 / a : a;
    for (unsigned k = n < 0 ? -n : n; k; k >>= 1) {
        if (k & 1) r *= b;
        b *= b;
    }
    return r;
}

// Padded layout of one member: element (x, y, z) at x * N (N + 1) + y * N + z.
__host__ __device__ constexpr size_t plane_stride(int N) { return (size_t)N * (N + 1); }
__host__ __device__ constexpr size_t member_stride(int N) { return (size_t)N * plane_stride(N); }

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
