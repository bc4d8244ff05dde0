// psc_ic_modes.hpp -- the visiting order of the initial modes and the
// Park-Miller stream at an arbitrary draw, shared by the host generator, the
// device generator and a CPU test (plain C++ when compiled without HIP).
//
// initialize() visits, per field (initialize.cpp:685-775; initialize_GW uses
// the same order, initialize.cpp:355-428, without the DC mode), with
// h = N/2 - 1:
//   for kz in {0, N/2}:                          the two real planes, half each,
//     A1  i in {0, N/2}, j = 1..h                conjugate (i, N-j, kz)
//     A2  j in {0, N/2}, i = 1..h                conjugate (N-i, j, kz)
//     A3  i = 1..h, j = 1..h then N/2+1..N-1     conjugate (N-i, N-j, kz)
//   B   kz = 1..h, i = 0..N-1, j = 0..N-1        the interior, no conjugate
//   C   kz, i, j in {0, N/2} (8 or 7 corners)    real modes
// and each mode takes 4 consecutive draws of rand_uniform.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/psc.h"

#include "psc_model.hpp"
// No contraction into fused multiply-adds inside the mode arithmetic, so that
// host and device round every product and sum alike.
#if defined(__clang__)
#define PSC_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define PSC_NO_CONTRACT
#endif

namespace psc_ic {

PSC_HD inline double pw2(double x) { return x * x; }

// Lattice sizes of the library: 2^k or 3 * 2^k, 16 <= N <= 512 (the lengths
// its transforms take; a size without launch shapes is refused at psc_create).
PSC_HD inline bool lattice_size_ok(int N) {
    const int m = N % 3 == 0 ? N / 3 : N;
    return N >= 16 && N <= 512 && N % 2 == 0 && (m & (m - 1)) == 0;
}

// Modes visited per field: 2 (4h + 2h^2) plane modes, h N^2 interior modes,
// 8 corners (7 without the DC mode).
PSC_HD inline uint64_t mode_count(int N, bool dc) {
    const uint64_t h = N / 2 - 1;
    return 2 * (4 * h + 2 * h * h) + h * (uint64_t)N * N + (dc ? 8 : 7);
}

struct Mode {
    uint64_t v;      // position in the visiting order (draws 4v .. 4v+3)
    int i, j, k;     // half-spectrum index, kz = k in [0, N/2]
    int ci, cj;      // index of the conjugate filled in the same plane, or -1
    bool real;       // corner mode: imaginary parts set to zero
};

// The mode handled by work item t in [0, mode_count).  Plane modes come first
// in visiting order; interior modes are ordered with kz fastest (so that
// neighbouring work items touch neighbouring kz); corners last.
PSC_HD inline Mode mode_at(int N, bool dc, uint64_t t) {
    const int h = N / 2 - 1;
    const uint64_t S = 4 * (uint64_t)h + 2 * (uint64_t)h * h;   // modes per real plane
    Mode m{};
    m.ci = m.cj = -1;
    m.real = false;
    if (t < 2 * S) {
        m.v = t;
        m.k = t < S ? 0 : N / 2;
        const uint64_t r = t % S;
        if (r < 2 * (uint64_t)h) {
            m.i = r < (uint64_t)h ? 0 : N / 2;
            m.j = 1 + (int)(r % h);
            m.ci = m.i;
            m.cj = N - m.j;
        } else if (r < 4 * (uint64_t)h) {
            const uint64_t q = r - 2 * h;
            m.j = q < (uint64_t)h ? 0 : N / 2;
            m.i = 1 + (int)(q % h);
            m.ci = N - m.i;
            m.cj = m.j;
        } else {
            const uint64_t q = r - 4 * h;
            m.i = 1 + (int)(q / (2 * h));
            const int p = (int)(q % (2 * h));
            m.j = p < h ? 1 + p : N / 2 + 1 + (p - h);
            m.ci = N - m.i;
            m.cj = N - m.j;
        }
        return m;
    }
    const uint64_t nb = (uint64_t)h * N * N;
    const uint64_t u = t - 2 * S;
    if (u < nb) {
        m.k = 1 + (int)(u % h);
        const uint64_t ij = u / h;
        m.i = (int)(ij / N);
        m.j = (int)(ij % N);
        m.v = 2 * S + ((uint64_t)(m.k - 1) * N + m.i) * N + m.j;
        return m;
    }
    int c = (int)(u - nb);
    m.v = 2 * S + nb + c;
    if (!dc) ++c;   // the DC corner is not visited
    m.k = (c >> 2) & 1 ? N / 2 : 0;
    m.i = (c >> 1) & 1 ? N / 2 : 0;
    m.j = c & 1 ? N / 2 : 0;
    m.real = true;
    return m;
}

// The low-pass of the initial modes (kcutoff, initialize.cpp:138-145 and
// 281-294): a mode with dp2 * s > k2cutoff, s = px^2 + py^2 + pz^2, is zero
// and takes no draws.  s is an integer (exact in double) and dp2 * s rounds
// monotonically in s, so the cut is s > smax for the largest s with
// dp2 * s <= k2cutoff, and the number of kept modes before any mode (its
// draw offset / 4) is integer counting: in closed form along a row, through
// a table of row offsets over the interior (the host fills it).  smax < 0:
// nothing is cut.
struct Cut {
    int64_t smax = -1;
    PSC_HD bool on() const { return smax >= 0; }
    PSC_HD bool keeps(int64_t s) const { return smax < 0 || s <= smax; }
    // integers y in [lo, hi] with y^2 <= R (all of them without a cut)
    PSC_HD int64_t count(int64_t R, int64_t lo, int64_t hi) const {
        if (hi < lo) return 0;
        if (smax < 0) return hi - lo + 1;
        if (R < 0) return 0;
        int64_t r = (int64_t)sqrt((double)R);
        while (r * r > R) --r;
        while ((r + 1) * (r + 1) <= R) ++r;
        const int64_t a = lo > -r ? lo : -r, b = hi < r ? hi : r;
        return b >= a ? b - a + 1 : 0;
    }
};

// k2cutoff = kcutoff^2 when kcutoff < 2 pi N / L (L = 2 pi / kIR), else no
// cut (initialize.cpp:139, 282); kcutoff <= 0 cuts nothing either.
inline Cut make_cut(const psc_config& c, double kcutoff) {
    const double pi = 3.14159265358979323846;
    const double L = 2.0 * pi / c.kIR;
    const double k2 = kcutoff < 2.0 * pi * (double)c.N / L ? kcutoff * kcutoff : 0.0;
    Cut cut;
    if (!(k2 > 0.0)) return cut;
    const double dp2 = c.kIR * c.kIR;
    const int64_t top = 3 * (int64_t)(c.N / 2) * (c.N / 2);
    int64_t s = (int64_t)(k2 / dp2);
    if (s > top) s = top;
    while (s < top && dp2 * (double)(s + 1) <= k2) ++s;
    while (s >= 0 && dp2 * (double)s > k2) --s;
    cut.smax = s < 0 ? 0 : s;   // s = 0 (the DC mode) always passes: 0 <= k2
    return cut;
}

// Kept modes among the first r visits of the real plane kz (r <= 4h + 2h^2,
// the plane part of mode_at's order).
PSC_HD inline uint64_t plane_kept_before(int N, const Cut& cut, int kz, uint64_t r) {
    const int64_t h = N / 2 - 1, K = (int64_t)kz * kz;
    auto clampi = [](int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
    uint64_t n = 0;
    for (int a = 0; a < 2; ++a) {   // i = 0, N/2; j = 1..h
        const int64_t ii = a ? N / 2 : 0;
        n += cut.count(cut.smax - ii * ii - K, 1, clampi((int64_t)r - a * h, 0, h));
    }
    if ((int64_t)r <= 2 * h) return n;
    for (int a = 0; a < 2; ++a) {   // j = 0, N/2; i = 1..h
        const int64_t jj = a ? N / 2 : 0;
        n += cut.count(cut.smax - jj * jj - K, 1, clampi((int64_t)r - 2 * h - a * h, 0, h));
    }
    if ((int64_t)r <= 4 * h) return n;
    const int64_t q = (int64_t)r - 4 * h, full = q / (2 * h), rem = q % (2 * h);
    for (int64_t i = 1; i <= full; ++i) {   // j = 1..h (py = j), then py = -h..-1
        const int64_t R = cut.smax - i * i - K;
        n += cut.count(R, 1, h) + cut.count(R, -h, -1);
    }
    if (rem > 0) {
        const int64_t i = full + 1, R = cut.smax - i * i - K;
        n += rem <= h ? cut.count(R, 1, rem) : cut.count(R, 1, h) + cut.count(R, -h, -h + (rem - h) - 1);
    }
    return n;
}

// Kept modes of the interior row (kz, i), and of its first j visits.
PSC_HD inline uint64_t row_kept_before(int N, const Cut& cut, int kz, int i, int j) {
    const int64_t px = i <= N / 2 ? i : i - N;
    const int64_t R = cut.smax - px * px - (int64_t)kz * kz;
    if (j <= N / 2 + 1) return cut.count(R, 0, j - 1);
    return cut.count(R, 0, N / 2) + cut.count(R, -(N / 2 - 1), (int64_t)j - 1 - N);
}

// The draw-offset / 4 of every mode (the number of kept modes visited before
// it) for mode_at's order: planes, interior rows through `rowpre`, corners.
struct KeptIndex {
    Cut cut;
    uint64_t plane0 = 0;     // kept modes of the kz = 0 plane
    uint64_t planes = 0;     // of both planes
    uint64_t interior = 0;   // of the interior
    const uint64_t* rowpre = nullptr;   // [(kz - 1) N + i]: kept interior modes before the row
    PSC_HD uint64_t rank(int N, bool dc, const Mode& m, uint64_t t) const {
        const int h = N / 2 - 1;
        const uint64_t S = 4 * (uint64_t)h + 2 * (uint64_t)h * h;
        if (t < 2 * S) return (t < S ? 0 : plane0) + plane_kept_before(N, cut, m.k, t % S);
        if (t < 2 * S + (uint64_t)h * N * N)
            return planes + rowpre[(size_t)(m.k - 1) * N + m.i] + row_kept_before(N, cut, m.k, m.i, m.j);
        uint64_t n = planes + interior;
        for (int c = dc ? 0 : 1; c < 8; ++c) {
            const int k = (c >> 2) & 1 ? N / 2 : 0, i = (c >> 1) & 1 ? N / 2 : 0, j = c & 1 ? N / 2 : 0;
            if (k == m.k && i == m.i && j == m.j) break;
            if (cut.keeps((int64_t)k * k + (int64_t)i * i + (int64_t)j * j)) ++n;
        }
        return n;
    }
};

// Fills the interior row table and the plane/interior totals; returns the
// number of kept modes in the whole visit (with or without the DC corner).
inline uint64_t build_kept_index(int N, bool dc, const Cut& cut, KeptIndex& k, std::vector<uint64_t>& rowpre) {
    const int h = N / 2 - 1;
    const uint64_t S = 4 * (uint64_t)h + 2 * (uint64_t)h * h;
    k.cut = cut;
    k.plane0 = plane_kept_before(N, cut, 0, S);
    k.planes = k.plane0 + plane_kept_before(N, cut, N / 2, S);
    rowpre.assign((size_t)h * N, 0);
    uint64_t acc = 0;
    for (int kz = 1; kz <= h; ++kz)
        for (int i = 0; i < N; ++i) {
            rowpre[(size_t)(kz - 1) * N + i] = acc;
            acc += row_kept_before(N, cut, kz, i, N);
        }
    k.interior = acc;
    k.rowpre = rowpre.data();
    uint64_t corners = 0;
    for (int c = dc ? 0 : 1; c < 8; ++c) {
        const int64_t v = N / 2;
        const int64_t s = ((c >> 2) & 1 ? v * v : 0) + ((c >> 1) & 1 ? v * v : 0) + (c & 1 ? v * v : 0);
        if (cut.keeps(s)) ++corners;
    }
    return k.planes + k.interior + corners;
}

// rand_uniform (initialize.cpp:13-28) positioned at draw d (0-based): the
// state after the 99-step warm-up and d + 1 advances is 16807^(100 + d) seed
// mod 2^31 - 1 (Schrage's step is exactly that multiplication).  Seeds below 1
// give the constant 0.33 (initialize.cpp:15).
struct Stream {
    static constexpr uint64_t M = 2147483647ull;
    uint64_t s;
    bool debug;
    PSC_HD static uint64_t mulmod(uint64_t a, uint64_t b) {
        uint64_t x = a * b;             // < 2^62
        x = (x & M) + (x >> 31);        // < 2^32
        x = (x & M) + (x >> 31);        // <= M + 1
        return x >= M ? x - M : x;
    }
    PSC_HD static Stream at(int seed, uint64_t d) {
        Stream r;
        r.debug = seed < 1;
        uint64_t e = 100 + d, base = 16807, acc = 1;
        while (e) {
            if (e & 1) acc = mulmod(acc, base);
            base = mulmod(base, base);
            e >>= 1;
        }
        r.s = mulmod(acc, r.debug ? 1u : (uint64_t)seed);
        return r;
    }
    // The current draw; next() moves to the following one.
    PSC_HD double uniform() const { return debug ? 0.33 : (double)s / 2147483647.0; }
    PSC_HD void next() { s = mulmod(s, 16807); }
};

// orthogonal_unit_vectors / TT_tensors (initialize.cpp:66-122): the
// polarisation tensor e = u u - v v + u v + v u of the mode with momentum k,
// u = k x ref / |k x ref|, v = k x u / |k x u| with ref = y for k along z and z
// otherwise; u and v each change sign once if any of their components is
// negative.
PSC_HD inline void polarisation(const double k[3], double e[3][3]) {
    PSC_NO_CONTRACT
    const double kn = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    const double h[3] = {k[0] / kn, k[1] / kn, k[2] / kn};
    const bool z = fabs(h[0]) < 1e-9 && fabs(h[1]) < 1e-9 && fabs(h[2] - 1.0) < 1e-9;
    const double ref[3] = {0.0, z ? 1.0 : 0.0, z ? 0.0 : 1.0};
    double u[3] = {h[1] * ref[2] - h[2] * ref[1], h[2] * ref[0] - h[0] * ref[2], h[0] * ref[1] - h[1] * ref[0]};
    const double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int i = 0; i < 3; ++i) u[i] /= un;
    double v[3] = {h[1] * u[2] - h[2] * u[1], h[2] * u[0] - h[0] * u[2], h[0] * u[1] - h[1] * u[0]};
    const double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) v[i] /= vn;
    for (int i = 0; i < 3; ++i)
        if (u[i] < 0) {
            for (int q = 0; q < 3; ++q) u[q] = -u[q];
            break;
        }
    for (int i = 0; i < 3; ++i)
        if (v[i] < 0) {
            for (int q = 0; q < 3; ++q) v[q] = -v[q];
            break;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) e[i][j] = u[i] * u[j] - v[i] * v[j] + u[i] * v[j] + v[i] * u[j];
}

// The potential terms (potential_energy, model.h:35-73) and the effective
// masses at the homogeneous values (effective_mass, model.h:96-98) of the
// build's model for one member's couplings.
struct Pot {
    int nflds;
    psc::Couplings<double> cp;
    Pot(int nf, const psc_model& m) : nflds(nf), cp(m) {}
    int terms() const { return psc::Model::num_potential_terms(nflds); }
    double term(int t, const double* f) const {
        PSC_NO_CONTRACT
        return psc::Model::potential_energy<double>(t, f, cp);
    }
    void meff(double* m, const double* f) const {
        PSC_NO_CONTRACT
        double ms[2];
        psc::Model::effective_mass<double>(ms, f, cp);
        m[0] = ms[0];
        if (nflds == 2) m[1] = ms[1];
    }
};

// The homogeneous background of initialize() (initialize.cpp:624-668): H from
// the Friedmann equation with the initial field values and derivatives, a,
// da/dt = H a, and the printed d2a/dt2; all zero but a without expansion.
inline void background(const psc_config& c, const Pot& V, const psc_ic_params& ic, double* a, double* ad,
                       double* addp, double* H) {
    PSC_NO_CONTRACT
    double f0[2] = {ic.f0[0], c.nflds > 1 ? ic.f0[1] : 0.0};
    double fdsq = 0.0, pot = 0.0;
    for (int q = 0; q < c.nflds; ++q) fdsq += pw2(ic.fd0[q]);
    for (int t = 0; t < V.terms(); ++t) pot += V.term(t, f0);
    double h = (fdsq / 2.0 + pot) * pw2(c.fStar / c.MPl) / 3.0;
    *H = std::sqrt(h);
    *a = ic.a0;
    *ad = *H * ic.a0;
    *addp = pw2(c.fStar / c.MPl) / 3.0 * ((c.alpha - 2) * fdsq + (c.alpha + 1) * pot) * ic.a0;
    if (c.expansion == PSC_EXPANSION_STATIC) *H = *ad = *addp = 0.0;
}

}  // namespace psc_ic
