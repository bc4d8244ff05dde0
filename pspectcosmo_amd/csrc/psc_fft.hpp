// psc_fft.hpp -- device-side Stockham FFT building blocks for gfx950.
//
// A line of N complex points is transformed cooperatively by TPL threads.
// Thread t keeps elements x[t + TPL*m] (m < N/TPL) in registers before the
// first stage and after the last stage, so global loads/stores, and the fused
// pointwise work around a transform, never touch LDS.  Between stages the line
// goes through LDS (in place, two barriers per stage).  Stage s with radix R
// and Ns = product of earlier radices follows the Stockham autosort recursion:
//   v[r] = x[j + r N/R] * W_{Ns R}^{r (j mod Ns)}, v = DFT_R(v),
//   x'[(j / Ns) Ns R + (j mod Ns) + r Ns] = v[r],
// for butterflies j < N/R; the result is in natural order.
#pragma once
#include <hip/hip_runtime.h>

namespace psc {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}

// cos / sin of 2 pi m / 64, m < 64 (constant-folded after unrolling).
struct Unit64 {
    static constexpr double c[64] = {
        1.0, 0.99518472667219688624, 0.98078528040323044913, 0.95694033573220886494,
        0.92387953251128675613, 0.88192126434835502971, 0.83146961230254523708, 0.77301045336273696081,
        0.70710678118654752440, 0.63439328416364549822, 0.55557023301960222474, 0.47139673682599764856,
        0.38268343236508977173, 0.29028467725446236764, 0.19509032201612826785, 0.098017140329560601994,
        0.0, -0.098017140329560601994, -0.19509032201612826785, -0.29028467725446236764,
        -0.38268343236508977173, -0.47139673682599764856, -0.55557023301960222474, -0.63439328416364549822,
        -0.70710678118654752440, -0.77301045336273696081, -0.83146961230254523708, -0.88192126434835502971,
        -0.92387953251128675613, -0.95694033573220886494, -0.98078528040323044913, -0.99518472667219688624,
        -1.0, -0.99518472667219688624, -0.98078528040323044913, -0.95694033573220886494,
        -0.92387953251128675613, -0.88192126434835502971, -0.83146961230254523708, -0.77301045336273696081,
        -0.70710678118654752440, -0.63439328416364549822, -0.55557023301960222474, -0.47139673682599764856,
        -0.38268343236508977173, -0.29028467725446236764, -0.19509032201612826785, -0.098017140329560601994,
        0.0, 0.098017140329560601994, 0.19509032201612826785, 0.29028467725446236764,
        0.38268343236508977173, 0.47139673682599764856, 0.55557023301960222474, 0.63439328416364549822,
        0.70710678118654752440, 0.77301045336273696081, 0.83146961230254523708, 0.88192126434835502971,
        0.92387953251128675613, 0.95694033573220886494, 0.98078528040323044913, 0.99518472667219688624};
    static constexpr double cosv(int m) { return c[m & 63]; }
    static constexpr double sinv(int m) { return c[(m + 48) & 63]; }  // sin x = cos(x - pi/2)
};

// v * exp(SIGN 2 pi i m / 64) for a compile-time m.
template <int SIGN, int M, typename T>
__device__ __forceinline__ cpx<T> rot64(cpx<T> v) {
    if constexpr (M == 0) {
        return v;
    } else if constexpr (M == 16) {
        return SIGN > 0 ? cpx<T>{-v.y, v.x} : cpx<T>{v.y, -v.x};
    } else {
        constexpr T c = (T)Unit64::cosv(M);
        constexpr T s = (T)(SIGN * Unit64::sinv(M));
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int bitrev(int i, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b)
        if (i & (1 << b)) r |= 1 << (bits - 1 - b);
    return r;
}

// In-register DFT of length R (power of two <= 64; 3 * 2^k through Dft3M) by a
// radix-2 DIT network:
// out[k] = sum_r in[r] exp(SIGN 2 pi i r k / R).
template <int R, int SIGN, typename T, int LEN = 2>
struct DftLevel {
    template <int I, int K>
    static __device__ __forceinline__ void bfly(cpx<T>* v) {
        if constexpr (I < R) {
            if constexpr (K < LEN / 2) {
                cpx<T> a = v[I + K];
                cpx<T> b = rot64<SIGN, K * (64 / LEN)>(v[I + K + LEN / 2]);
                v[I + K] = cadd(a, b);
                v[I + K + LEN / 2] = csub(a, b);
                bfly<I, K + 1>(v);
            } else {
                bfly<I + LEN, 0>(v);
            }
        }
    }
    static __device__ __forceinline__ void run(cpx<T>* v) {
        bfly<0, 0>(v);
        if constexpr (LEN < R) DftLevel<R, SIGN, T, LEN * 2>::run(v);
    }
};

template <int R, int I = 0>
struct BitRevPermute {
    template <typename T>
    static __device__ __forceinline__ void run(cpx<T>* v) {
        if constexpr (I < R) {
            constexpr int J = bitrev(I, ilog2(R));
            if constexpr (J > I) {
                cpx<T> tmp = v[I];
                v[I] = v[J];
                v[J] = tmp;
            }
            BitRevPermute<R, I + 1>::run(v);
        }
    }
};

// cos(2 pi m / n) and sin(2 pi m / n) at compile time (Taylor series after
// reduction to [-pi/4, pi/4]; exact at the multiples of pi/4), for the
// twiddles of the radix-3 lengths.
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
// The angle 2 pi m / n is 2 pi (8 m mod 8n) / (8n): octant o = floor(8 m / n)
// and remainder, so that the series argument stays within pi/4.
constexpr double unit_cos(long m, long n) {
    m = ((m % n) + n) % n;
    const long o = (8 * m) / n;                          // octant 0..7
    const double r = kTwoPi * (double)(8 * m - o * n) / (8.0 * (double)n);   // in [0, pi/4)
    switch (o) {
        case 0: return taylor_cos(r);
        case 1: return taylor_sin(kTwoPi / 8 - r);
        case 2: return -taylor_sin(r);
        case 3: return -taylor_cos(kTwoPi / 8 - r);
        case 4: return -taylor_cos(r);
        case 5: return -taylor_sin(kTwoPi / 8 - r);
        case 6: return taylor_sin(r);
        default: return taylor_cos(kTwoPi / 8 - r);
    }
}
constexpr double unit_sin(long m, long n) { return unit_cos(4 * m - n, 4 * n); }   // sin x = cos(x - pi/2)

// v * exp(SIGN 2 pi i M / R) for compile-time M, R.
template <int SIGN, int M, int R, typename T>
__device__ __forceinline__ cpx<T> rot_n(cpx<T> v) {
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) {
        return v;
    } else if constexpr (4 * m == R) {
        return SIGN > 0 ? cpx<T>{-v.y, v.x} : cpx<T>{v.y, -v.x};
    } else if constexpr (2 * m == R) {
        return cpx<T>{-v.x, -v.y};
    } else if constexpr (4 * m == 3 * R) {
        return SIGN > 0 ? cpx<T>{v.y, -v.x} : cpx<T>{-v.y, v.x};
    } else {
        constexpr T c = (T)unit_cos(m, R);
        constexpr T s = (T)(SIGN * unit_sin(m, R));
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cpx<T>* v);

// In-register DFT of length R = 3 M (M a power of two): the three
// interleaved sub-sequences x[r0 + 3 r1] are transformed with length M, each
// output k1 of sub-sequence r0 is turned by exp(SIGN 2 pi i r0 k1 / R), and a
// length-3 DFT over r0 gives the outputs k1, k1 + M, k1 + 2M.
template <int R, int SIGN, typename T>
struct Dft3M {
    static constexpr int M = R / 3;
    template <int K>
    static __device__ __forceinline__ void combine(cpx<T> (&y)[3][M], cpx<T>* v) {
        if constexpr (K < M) {
            const cpx<T> a = y[0][K];
            const cpx<T> b = rot_n<SIGN, K, R>(y[1][K]);
            const cpx<T> c = rot_n<SIGN, 2 * K, R>(y[2][K]);
            const cpx<T> s = cadd(b, c), d = csub(b, c);
            constexpr T h = (T)0.86602540378443864676;   // sqrt(3) / 2
            const cpx<T> m = {a.x - (T)0.5 * s.x, a.y - (T)0.5 * s.y};
            // SIGN sqrt(3)/2 i d
            const cpx<T> u = SIGN > 0 ? cpx<T>{-h * d.y, h * d.x} : cpx<T>{h * d.y, -h * d.x};
            v[K] = cadd(a, s);
            v[K + M] = cadd(m, u);
            v[K + 2 * M] = csub(m, u);
            combine<K + 1>(y, v);
        }
    }
    static __device__ __forceinline__ void run(cpx<T>* v) {
        cpx<T> y[3][M];
#pragma unroll
        for (int r0 = 0; r0 < 3; ++r0)
#pragma unroll
            for (int r1 = 0; r1 < M; ++r1) y[r0][r1] = v[r0 + 3 * r1];
#pragma unroll
        for (int r0 = 0; r0 < 3; ++r0) dft<M, SIGN>(y[r0]);
        combine<0>(y, v);
    }
};

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cpx<T>* v) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (is_pow2(R)) {
        BitRevPermute<R>::run(v);
        DftLevel<R, SIGN, T, 2>::run(v);
    } else {
        static_assert(R % 3 == 0 && is_pow2(R / 3), "radix must be 2^k or 3 * 2^k");
        Dft3M<R, SIGN, T>::run(v);
    }
}

template <int... Rs>
struct Radices {};

// Radix plans.  With first radix R1 a line is handled by TPL = N / R1 threads:
// one butterfly per thread in the first stage; later stages (radix <= R1) give
// each thread N / (R * TPL) butterflies.

// LDS layout for lines whose TPL threads are adjacent lanes (contiguous-axis
// pass): line c, element n -> c (N + N/P) + n + n/P (one pad slot every P
// elements breaks the stride-R1 bank pattern of the first stage's writes).
template <int N, int P>
struct LayoutRow {
    static constexpr int stride = N + N / P;
    __device__ __forceinline__ static int at(int c, int n) { return c * stride + n + n / P; }
    static constexpr int size(int lines) { return lines * stride; }
};
// LDS layout for lines spread across lanes (strided-axis pass): consecutive
// lanes hold consecutive columns c, element n -> n * TC + c.
template <int N, int TC>
struct LayoutCol {
    __device__ __forceinline__ static int at(int c, int n) { return n * TC + c; }
    static constexpr int size(int) { return N * TC; }
};

// One Stockham stage.
template <typename T, int N, int TPL, int SIGN, int R, int NS, bool FIRST, bool LAST, class Lay>
__device__ __forceinline__ void fft_stage(cpx<T>* e, int t, int c, cpx<T>* lds,
                                          const cpx<T>* __restrict__ tw) {
    constexpr int Q = N / (R * TPL);
    static_assert(Q >= 1, "radix too large for the thread count");
    static_assert(N % (R * TPL) == 0, "every thread must take a whole number of butterflies per stage "
                                      "(the later radices must divide the first)");
    cpx<T> v[Q][R];
    if constexpr (FIRST) {
        static_assert(Q == 1, "first stage must have one butterfly per thread");
#pragma unroll
        for (int r = 0; r < R; ++r) v[0][r] = e[r];
    } else {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int j = t + TPL * q;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = lds[Lay::at(c, j + r * (N / R))];
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + TPL * q;
        if constexpr (NS > 1) {
            const int m = (j % NS) * (N / (NS * R));
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cpx<T> w = tw[r * m];
                if (SIGN > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
        }
        dft<R, SIGN>(v[q]);
    }
    if constexpr (LAST) {
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) e[q + r * Q] = v[q][r];
    } else {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int j = t + TPL * q;
            const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[Lay::at(c, base + r * NS)] = v[q][r];
        }
    }
}

template <typename T, int N, int TPL, int SIGN, class Lay, int NS, bool FIRST, int... Rs>
struct StageChain;
template <typename T, int N, int TPL, int SIGN, class Lay, int NS, bool FIRST, int R>
struct StageChain<T, N, TPL, SIGN, Lay, NS, FIRST, R> {
    static __device__ __forceinline__ void run(cpx<T>* e, int t, int c, cpx<T>* lds,
                                               const cpx<T>* tw) {
        fft_stage<T, N, TPL, SIGN, R, NS, FIRST, true, Lay>(e, t, c, lds, tw);
    }
};
template <typename T, int N, int TPL, int SIGN, class Lay, int NS, bool FIRST, int R, int R2,
          int... Rest>
struct StageChain<T, N, TPL, SIGN, Lay, NS, FIRST, R, R2, Rest...> {
    static __device__ __forceinline__ void run(cpx<T>* e, int t, int c, cpx<T>* lds,
                                               const cpx<T>* tw) {
        fft_stage<T, N, TPL, SIGN, R, NS, FIRST, false, Lay>(e, t, c, lds, tw);
        StageChain<T, N, TPL, SIGN, Lay, NS * R, false, R2, Rest...>::run(e, t, c, lds, tw);
    }
};

template <typename T, int N, int TPL, int SIGN, class Lay, class Rad>
struct LineFftImpl;
template <typename T, int N, int TPL, int SIGN, class Lay, int... Rs>
struct LineFftImpl<T, N, TPL, SIGN, Lay, Radices<Rs...>> {
    static __device__ __forceinline__ void run(cpx<T>* e, int t, int c, cpx<T>* lds,
                                               const cpx<T>* tw) {
        StageChain<T, N, TPL, SIGN, Lay, 1, true, Rs...>::run(e, t, c, lds, tw);
    }
};

template <int R, int... Rs>
constexpr int first_radix(Radices<R, Rs...>) { return R; }

// Transform the line held as e[m] = x[t + TPL m], TPL = N / first radix;
// result in the same layout.  SIGN = -1: forward (e^{-ikx}); +1: inverse,
// unnormalised.  tw[m] = exp(-2 pi i m / N).
template <typename T, int N, class Rad, int SIGN, class Lay>
__device__ __forceinline__ void line_fft(cpx<T>* e, int t, int c, cpx<T>* lds,
                                         const cpx<T>* __restrict__ tw) {
    constexpr int TPL = N / first_radix(Rad{});
    LineFftImpl<T, N, TPL, SIGN, Lay, Rad>::run(e, t, c, lds, tw);
}

}  // namespace psc
