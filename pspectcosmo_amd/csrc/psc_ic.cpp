// psc_ic.cpp -- host-side initial conditions of the product library.
//
// Restates initialize() (initialize.cpp:562-810) for the non-inflation
// Rayleigh branch (set_mode, initialize.cpp:128-220) and initialize_GW
// (initialize.cpp:341-470), writing FFTW-layout FP64 half spectra.  The output
// is loaded onto the device with psc_load_ic_k / psc_load_gw_k.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include "../../include/psc.h"
#include "psc_ic_modes.hpp"

// Bit-exact agreement with other IEEE restatements of initialize(): no FMA
// contraction in this file, and the phase's sine and cosine from one glibc
// sincos call (what g++ -O3 emits for the reference's separate cos/sin calls at
// initialize.cpp:185-191).
#pragma clang fp contract(off)

namespace {

const double kPi = 3.14159265358979323846264338327950288;
inline double pw2(double x) { return x * x; }
inline int sidx(int i, int N) { return i <= N / 2 ? i : i - N; }

// rand_uniform (initialize.cpp:13-28): Park-Miller minimal standard with
// Schrage's factorisation; the first call advances 99 extra times.
class Rng {
  public:
    explicit Rng(int seed) : seed_(seed), next_(seed > 0 ? seed : 1) {}
    // Position the stream so the next uniform() returns draw number d
    // (0-based): state = 16807^(99 + d + 1 - 1) x seed mod (2^31 - 1) before
    // the advance inside uniform(), i.e. skip-ahead by modular exponentiation.
    static Rng at(int seed, uint64_t d) {
        Rng r(seed);
        r.warm_ = true;
        uint64_t e = 99 + d, base = 16807, acc = 1;
        const uint64_t m = 2147483647ull;
        while (e) {
            if (e & 1) acc = acc * base % m;
            base = base * base % m;
            e >>= 1;
        }
        r.next_ = (int)(acc * (uint64_t)r.next_ % m);
        return r;
    }
    double uniform() {
        if (seed_ < 1) return 0.33;
        if (!warm_) {
            for (int i = 1; i < 100; ++i) step();
            warm_ = true;
        }
        step();
        return (double)next_ / 2147483647.0;
    }

  private:
    void step() {
        next_ = 16807 * (next_ % 127773) - 2836 * (next_ / 127773);
        if (next_ < 0) next_ += 2147483647;
    }
    int seed_;
    int next_;
    bool warm_ = false;
};


using psc_ic::Pot;
using psc_ic::background;

// Visit the modes in that order; each mode below the cutoff consumes 4 draws
// from the stream `rng` handed to fn(rng, i, j, k, conjugate index or -1,
// isReal), and a cut mode is skipped (no call, no draws).  The stream starts
// at draw `first`.  The interior planes (the bulk) run on worker threads,
// each (kz, i) row starting from its own skip-ahead position (the kept modes
// before it, from `kept`), which leaves every mode with exactly the draws of
// the sequential order.
template <class F>
void for_each_mode(int N, bool dc, int seed, uint64_t first, const psc_ic::KeptIndex& kept, F&& fn) {
    const int Nh = N / 2 + 1;
    auto at = [&](int i, int j, int k) { return ((long)i * N + j) * Nh + k; };
    auto visit = [&](Rng& r, int i, int j, int k, long conj, bool real) {
        const int64_t px = sidx(i, N), py = sidx(j, N);
        if (kept.cut.keeps(px * px + py * py + (int64_t)k * k)) fn(r, i, j, k, conj, real);
    };
    Rng rng = Rng::at(seed, first);
    for (int k = 0; k <= N / 2; k += N / 2) {
        for (int i = 0; i <= N / 2; i += N / 2)
            for (int j = 1; j < N / 2; ++j) visit(rng, i, j, k, at(i, N - j, k), false);
        for (int j = 0; j <= N / 2; j += N / 2)
            for (int i = 1; i < N / 2; ++i) visit(rng, i, j, k, at(N - i, j, k), false);
        for (int i = 1; i < N / 2; ++i) {
            for (int j = 1; j < N / 2; ++j) visit(rng, i, j, k, at(N - i, N - j, k), false);
            for (int j = N / 2 + 1; j < N; ++j) visit(rng, i, j, k, at(N - i, N - j, k), false);
        }
    }
    const long rows = (long)(N / 2 - 1) * N;  // (kz, i) rows of N modes each
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (rows < 256) nt = 1;
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; ++w)
        pool.emplace_back([&, w] {
            for (long r = w; r < rows; r += nt) {
                const int k = 1 + (int)(r / N), i = (int)(r % N);
                Rng rr = Rng::at(seed, first + 4 * (kept.planes + kept.rowpre[r]));
                for (int j = 0; j < N; ++j) visit(rr, i, j, k, -1L, false);
            }
        });
    for (auto& t : pool) t.join();
    Rng rc = Rng::at(seed, first + 4 * (kept.planes + kept.interior));
    for (int k = 0; k <= N / 2; k += N / 2)
        for (int i = 0; i <= N / 2; i += N / 2)
            for (int j = 0; j <= N / 2; j += N / 2) {
                if (!dc && i == 0 && j == 0 && k == 0) continue;
                visit(rc, i, j, k, -1L, true);
            }
}

bool valid(const psc_config* c) {
    return c && psc_ic::lattice_size_ok(c->N) && c->nflds >= 1 &&
           c->nflds <= 2 && c->kIR > 0 && c->fStar > 0 && c->MPl > 0;
}


}  // namespace

extern "C" int psc_generate_ic(const psc_config* cfg, const psc_model* model, const psc_ic_params* ic,
                               double* f_k, double* pi_k, double* scalars_out) {
    if (!valid(cfg) || !model || !ic || !f_k || !pi_k || ic->a0 <= 0) return PSC_ERR_ARG;
    const int N = cfg->N, Nh = N / 2 + 1;
    const size_t nk = (size_t)N * N * Nh;
    const Pot V(cfg->nflds, *model);
    double a, ad, addp, H;
    background(*cfg, V, *ic, &a, &ad, &addp, &H);
    if (!(H >= 0.0)) return PSC_ERR_ARG;  // initialize.cpp:661-664
    const double L = 2.0 * kPi / cfg->kIR, dx = L / N;
    // Vacuum normalisation (omegaStar/fStar) (N/dx)^{3/2} / sqrt(2): the
    // commented DeltaF_DeltaPi of parameters.h:132.
    const double norm = cfg->omegaStar / cfg->fStar * std::pow(N / dx, 1.5) / std::sqrt(2.0);
    const double f0[2] = {ic->f0[0], cfg->nflds > 1 ? ic->f0[1] : 0.0};
    double meff[2] = {0, 0};
    V.meff(meff, f0);
    const double dp2 = pw2(cfg->kIR), ap = std::pow(a, cfg->alpha - 3);
    const double N3 = (double)N * N * N;
    psc_ic::KeptIndex kept;
    std::vector<uint64_t> rowpre;
    const uint64_t M = psc_ic::build_kept_index(N, true, psc_ic::make_cut(*cfg, ic->kcutoff), kept, rowpre);
    for (int q = 0; q < cfg->nflds; ++q) {
        double* fq = f_k + 2 * nk * q;
        double* pq = pi_k + 2 * nk * q;
        std::memset(fq, 0, sizeof(double) * 2 * nk);   // the cut modes stay zero
        std::memset(pq, 0, sizeof(double) * 2 * nk);
        // four draws per kept mode (initialize.cpp:181-190), fields in sequence
        for_each_mode(N, true, ic->seed, 4 * (uint64_t)q * M, kept, [&](Rng& rng, int i, int j, int k, long conj, bool real) {
            const int px = sidx(i, N), py = sidx(j, N);
            const double p2 = dp2 * (pw2(px) + pw2(py) + pw2(k));
            const double om = (p2 + meff[q] > 0.0) ? std::sqrt(p2 + pw2(a) * meff[q]) : std::sqrt(p2);
            const double rms = om > 0.0 ? norm / std::sqrt(om) : 0.0;
            const double al = rms / std::sqrt(2.0) * std::sqrt(std::log(1.0 / rng.uniform())) / a;
            const double ar = rms / std::sqrt(2.0) * std::sqrt(std::log(1.0 / rng.uniform())) / a;
            double sn, cs;
            ::sincos(2.0 * kPi * rng.uniform(), &sn, &cs);
            const double rl = al * cs, il = al * sn;
            ::sincos(2.0 * kPi * rng.uniform(), &sn, &cs);
            const double rr = ar * cs, ir = ar * sn;
            double fr = rl + rr, fi = il + ir;
            double pr = ap * (-om * (il - ir) - H * fr);
            double pim = ap * (om * (rl - rr) - H * fi);
            if (real) fi = pim = 0.0;
            const long at = ((long)i * N + j) * Nh + k;
            fq[2 * at] = fr; fq[2 * at + 1] = fi;
            pq[2 * at] = pr; pq[2 * at + 1] = pim;
            if (conj >= 0) {
                fq[2 * conj] = fr; fq[2 * conj + 1] = -fi;
                pq[2 * conj] = pr; pq[2 * conj + 1] = -pim;
            }
        });
        fq[0] = ic->f0[q] * N3;
        fq[1] = 0.0;
        pq[0] = ap * ic->fd0[q] * N3;
        pq[1] = 0.0;
    }
    if (scalars_out) {
        scalars_out[0] = a;
        scalars_out[1] = ad;
        scalars_out[2] = addp;
        scalars_out[3] = H;
    }
    return PSC_OK;
}

extern "C" int psc_generate_gw_ic(const psc_config* cfg, const psc_model* model,
                                  const psc_ic_params* ic, double* u_k, double* pi_u_k) {
    if (!valid(cfg) || !model || !ic || !u_k || !pi_u_k || ic->a0 <= 0) return PSC_ERR_ARG;
    static const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
    const int N = cfg->N, Nh = N / 2 + 1;
    const size_t nk = (size_t)N * N * Nh;
    const Pot V(cfg->nflds, *model);
    double a, ad, addp, H;
    background(*cfg, V, *ic, &a, &ad, &addp, &H);
    // The GW draws follow the 4 draws of every kept scalar mode of every field.
    const psc_ic::Cut cut = psc_ic::make_cut(*cfg, ic->kcutoff);
    psc_ic::KeptIndex kept_f, kept;
    std::vector<uint64_t> rowpre_f, rowpre;
    const uint64_t skip = 4 * psc_ic::build_kept_index(N, true, cut, kept_f, rowpre_f) * (uint64_t)cfg->nflds;
    psc_ic::build_kept_index(N, false, cut, kept, rowpre);
    const double L = 2.0 * kPi / cfg->kIR, dx = L / N;
    const double norm = cfg->MPl * cfg->omegaStar / pw2(cfg->fStar) * std::pow(N / dx, 1.5);
    const double ap = std::pow(a, cfg->alpha - 3);
    std::memset(u_k, 0, sizeof(double) * 2 * nk * 6);
    std::memset(pi_u_k, 0, sizeof(double) * 2 * nk * 6);
    for_each_mode(N, false, ic->seed, skip, kept, [&](Rng& rng, int i, int j, int k, long conj, bool real) {
        const int px = sidx(i, N), py = sidx(j, N);
        const double kv[3] = {px * cfg->kIR, py * cfg->kIR, k * cfg->kIR};
        double e[3][3];
        psc_ic::polarisation(kv, e);
        const double kk = std::sqrt(pw2(cfg->kIR) * (pw2(px) + pw2(py) + pw2(k)));
        const double rms = norm / std::sqrt(kk);
        const double al = rms / std::sqrt(2.0) * std::sqrt(std::log(1.0 / rng.uniform())) / a;
        const double ar = rms / std::sqrt(2.0) * std::sqrt(std::log(1.0 / rng.uniform())) / a;
        double sn, cs;
        ::sincos(2.0 * kPi * rng.uniform(), &sn, &cs);
        const double rl = al * cs, il = al * sn;
        ::sincos(2.0 * kPi * rng.uniform(), &sn, &cs);
        const double rr = ar * cs, ir = ar * sn;
        const long at = ((long)i * N + j) * Nh + k;
        for (int c = 0; c < 6; ++c) {
            const double ec = e[ci[c]][cj[c]];
            const double ur = ec * (rl + rr), ui0 = ec * (il + ir);
            // initialize.cpp:317-318 multiplies the Hubble term by e_ij again.
            const double pr = ap * (ec * (-kk * (il - ir) - H * ur));
            const double pi0 = ap * (ec * (kk * (rl - rr) - H * ui0));
            const double ui = real ? 0.0 : ui0, pim = real ? 0.0 : pi0;
            double* uc = u_k + 2 * nk * c;
            double* pc = pi_u_k + 2 * nk * c;
            uc[2 * at] = ur; uc[2 * at + 1] = ui;
            pc[2 * at] = pr; pc[2 * at + 1] = pim;
            if (conj >= 0) {
                uc[2 * conj] = ur; uc[2 * conj + 1] = -ui;
                pc[2 * conj] = pr; pc[2 * conj + 1] = -pim;
            }
        }
    });
    return PSC_OK;
}
