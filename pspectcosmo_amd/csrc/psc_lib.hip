// psc_lib.hip -- host side of the HIP library: device state, kernel
// sequencing of the time step (evolve_VV_k_nopad, evolution.cpp:1131-1198),
// hipGraph capture, per-kernel timing and the C ABI of include/psc.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>
#include <memory>
#include <string>
#include <type_traits>
#include <type_traits>
#include <vector>

#include "../../include/psc.h"
#include "psc_kernels.hpp"
#include "psc_adjoint.hpp"
#include "psc_ic_device.hpp"
#include "psc_padded.hpp"

using namespace psc;

namespace {

enum Kind {
    K_Z_INV_DRIFT = 0,
    K_Y_INV,
    K_X_FORCE,
    K_Y_FWD,
    K_Z_FWD_KICK,
    K_SCALARS,
    K_Z_INV,
    K_Z_FWD_KICK_NR,
    K_DIAG,
    K_Z_INV_GW,
    K_X_FORCE_GW,
    K_Z_KICK_GW,
    K_Y_INV_GW,
    K_PAD_DRIFT,
    K_PAD_Z_INV,
    K_PAD_Y_INV,
    K_PAD_X_FORCE,
    K_PAD_Y_FWD,
    K_PAD_Z_FWD,
    K_PAD_KICK,
    K_NKINDS
};
const char* kKindNames[K_NKINDS] = {"z_inverse_drift", "y_inverse", "x_force", "y_forward",
                                    "z_forward_kick", "scalars", "z_inverse", "z_forward_kick_first",
                                    "diagnostics", "z_inverse_drift_gw", "x_source_gw", "z_forward_kick_gw",
                                    "y_inverse_gw", "pad_drift", "pad_z_inverse", "pad_y_inverse", "pad_x_force",
                                    "pad_y_forward", "pad_z_forward", "pad_kick"};

// Chunks of the first reduction level for passes with many blocks.
constexpr int kPreChunks = 512;

struct Timing {
    double ms = 0.0;
    long long count = 0;
    double bytes = 0.0;
};

struct Pending {
    int kind;
    hipEvent_t a, b;
};

struct EngineBase;

// The dealiased stepper's force chain on the 2N grid (defined below).
struct PadBase {
    virtual ~PadBase() {}
    // Drift (optional), padded V', extraction and both kicks; as force_chain.
    virtual int force_chain(psc_ctx* c, bool drift, bool reduce, double c1, double c2, int use_a_start) = 0;
    virtual int nV() const = 0;    // partial sums of V per force evaluation
    virtual int nKG() const = 0;   // partial sums of K, G per kick
    virtual size_t padded_elems() const = 0;
    // <V> over the padded lattice of the current state (kernel_a_k_pad's V_E).
    virtual int mean_V(psc_ctx* c, double* V) = 0;
};

// Device workspace of the adjoint, kept between calls with the same checkpoint
// spacing: checkpoints (f, pi, scalar state) at every segment start, one
// segment's recording (f, pi and the real-space fields at each substep
// boundary, pi_half per substep, the scalar state after each scalar update),
// the adjoint arrays and, when steps are graph-captured, the captured segment.
struct AdjWork {
    int K = 0, S = 0, nseg = 0;
    bool gw = false;                  // GW objective: u, pi_u carried
    std::vector<void*> ck;            // f, pi per segment
    MemberState* ck_ms = nullptr;
    std::vector<void*> f, p, r, ph;
    MemberState* rms = nullptr;
    void* fbar = nullptr;
    void* pbar = nullptr;
    psc::AdjState* st = nullptr;
    psc::AdjCoef* co = nullptr;
    double* p1 = nullptr;
    double* p3 = nullptr;
    double* p5 = nullptr;
    // GW: three packed arrays per entry, MS elements apart
    std::vector<void*> ck_u, ck_pu;   // per segment
    std::vector<void*> u, g, puh;     // u and real-space gradients per boundary, pi_u half per substep
    void* ubar = nullptr;
    void* pubar = nullptr;
    double* p1g = nullptr;
    double* p3g = nullptr;
    double* p5g = nullptr;
    double* pobj = nullptr;           // partials of the GW objective
    hipGraphExec_t g_rec = nullptr, g_bwd = nullptr;   // full segment: recording replay, sweep
    double bytes = 0.0;
    std::vector<void*> owned;
    ~AdjWork() {
        if (g_rec) (void)hipGraphExecDestroy(g_rec);
        if (g_bwd) (void)hipGraphExecDestroy(g_bwd);
        for (void* q : owned) (void)hipFree(q);
    }
};

}  // namespace

struct psc_ctx {
    psc_config cfg{};
    int N = 0, B = 1;
    size_t N3 = 0, es = 0;
    size_t MS = 0;   // member stride of the padded device arrays (elements)
    hipStream_t stream = nullptr;
    void* Pf = nullptr;
    void* Pp = nullptr;
    void* W = nullptr;   // [1 or 4][B][N^3]: fields / V', then 3 GW gradients / sources
    void* tw = nullptr;
    void* U = nullptr;   // GW: packed u pairs [3][B][N^3]
    void* PU = nullptr;  // GW: packed pi_u pairs [3][B][N^3]
    // dealiased stepper (cfg.pad == 2): padded work array, the kept padded
    // modes R[(N+1)^3], ghosts of the kz = 0, N/2 planes [2][N][N] of f and
    // pi, the scaled k = 0 mode, and twiddles of the 2N transforms
    void* Wp = nullptr;
    void* Rp = nullptr;
    void* Gf = nullptr;
    void* Gp = nullptr;
    void* dcp = nullptr;
    void* twp = nullptr;
    MemberState* ms = nullptr;
    double* partV = nullptr;
    double* partKG = nullptr;
    double* partV2 = nullptr;   // first-level reduction of partV (large N)
    double* partKG2 = nullptr;  // first-level reduction of partKG (large N)
    double* part = nullptr;   // diagnostics partials
    double* sums = nullptr;   // diagnostics totals
    size_t part_len = 0, sums_len = 0;
    std::vector<double> w;
    std::vector<psc_model> models;
    std::vector<char> loaded;
    int loaded_gw = 0;
    std::string err;
    EngineBase* eng = nullptr;
    PadBase* padeng = nullptr;
    int timing = 0;
    Timing kt[K_NKINDS];
    std::vector<hipEvent_t> ev_free;
    std::vector<Pending> pending;
    hipGraphExec_t g_mid = nullptr, g_last = nullptr;
    std::unique_ptr<AdjWork> adj;
    double dev_bytes = 0.0;
    int shape = 0;   // index into kShapes
    int pad_shape = -1;   // dealiased stepper: index into kShapes of its 2N shape
    int member_group = 0;   // members per group of the force chain (0: all)
    std::string report;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
};

#define HIPCHK(c, expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return (c)->fail(PSC_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

StepConsts consts_of(const psc_ctx* c) {
    StepConsts s{};
    s.dt = c->cfg.dt;
    s.kIR = c->cfg.kIR;
    s.f2m = (c->cfg.fStar / c->cfg.MPl) * (c->cfg.fStar / c->cfg.MPl);
    s.N3 = (double)c->N3;
    s.alpha = c->cfg.alpha;
    s.expansion_static = c->cfg.expansion == PSC_EXPANSION_STATIC;
    s.nflds = c->cfg.nflds;
    s.NV = c->padeng ? 8.0 * (double)c->N3 : (double)c->N3;   // lattice sites V is averaged over
    return s;
}

// Event-bracketed launch when timing is on.
template <class F>
int timed(psc_ctx* c, int kind, double bytes, F&& launch) {
    if (!c->timing) {
        launch();
        return PSC_OK;
    }
    hipEvent_t a, b;
    for (int q = 0; q < 2; ++q) {
        hipEvent_t e;
        if (!c->ev_free.empty()) {
            e = c->ev_free.back();
            c->ev_free.pop_back();
        } else {
            HIPCHK(c, hipEventCreate(&e));
        }
        (q == 0 ? a : b) = e;
    }
    HIPCHK(c, hipEventRecord(a, c->stream));
    launch();
    HIPCHK(c, hipEventRecord(b, c->stream));
    c->pending.push_back({kind, a, b});
    c->kt[kind].bytes = bytes;
    return PSC_OK;
}

int drain_timing(psc_ctx* c) {
    if (c->pending.empty()) return PSC_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& p : c->pending) {
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, p.a, p.b));
        c->kt[p.kind].ms += ms;
        c->kt[p.kind].count += 1;
        c->ev_free.push_back(p.a);
        c->ev_free.push_back(p.b);
    }
    c->pending.clear();
    return PSC_OK;
}

struct EngineBase {
    virtual ~EngineBase() = default;
    virtual int step(psc_ctx* c, int nsteps) = 0;
    // Real-space IFFT of src (Pf or Pp of one member) into W[member] as
    // normalised (phi, chi), with diagnostic sums into c->sums[0..5].
    virtual int real_space(psc_ctx* c, int member, const void* src) = 0;
    // Per-field K, G raw sums into c->sums[0..3].
    virtual int field_sums(psc_ctx* c, int member) = 0;
    // Binned spectrum partial sums into c->sums[0 .. 4*maxnumbins).
    virtual int spectra(psc_ctx* c, int member, int fld) = 0;
    // GW spectrum partial sums into c->sums[0 .. 3*maxnumbins + nlog).
    virtual int gw_spectra(psc_ctx* c, int member, int projection, int nlog) = 0;
    // Discrete-adjoint gradient (see psc_adjoint).
    virtual int adjoint(psc_ctx* c, const psc_adjoint_spec* sp, double* out, double* dchi) = 0;
    // Forward 3D transform of the packed real-space lattice in W[member] into dst[member].
    virtual int to_spectrum(psc_ctx* c, int member, void* dst) = 0;
    // Snapshot energy density `kind` (0 potential, 1 kinetic, 2 gradient) of one
    // member into the dense N^3 device array D; scale as k_x_density takes it.
    virtual int density(psc_ctx* c, int member, int kind, double* D, double scale) = 0;
};

// G drives every pass but the y-axis ones, which run on GY: a plan with more
// columns per block can suit the strided y passes where it would starve the
// x passes (their rows hold fused force, source and reduction work).
template <class G, class GY = G>
struct Engine final : EngineBase {
    using T = typename G::T;
    static constexpr int N = G::N;
    static int nV_blocks(const psc_ctx*) { return G::BLOCKS_COL; }
    static cpx<T>* P(void* p) { return static_cast<cpx<T>*>(p); }

    int launch_z_inverse(psc_ctx* c, bool drift, const cpx<T>* src, cpx<T>* Pf, const cpx<T>* Pp,
                         cpx<T>* W, const MemberState* ms, int members) {
        const double X = (double)members * G::N3 * sizeof(cpx<T>);
        dim3 grid(G::BLOCKS_ROW, members), block(G::THREADS_ROW);
        if (drift)
            return timed(c, K_Z_INV_DRIFT, 4 * X, [&] {
                if (streams(members))
                    hipLaunchKernelGGL((k_z_inverse<G, true, false, true, true>), grid, block, 0, c->stream, src, Pf,
                                       Pp, W, ms, P(c->tw), nullptr, nullptr);
                else
                    hipLaunchKernelGGL((k_z_inverse<G, true, false, true>), grid, block, 0, c->stream, src, Pf, Pp,
                                       W, ms, P(c->tw), nullptr, nullptr);
            });
        return timed(c, K_Z_INV, 2 * X, [&] {
            hipLaunchKernelGGL((k_z_inverse<G, false, false, true>), grid, block, 0, c->stream, src, Pf, Pp, W, ms,
                               P(c->tw), nullptr, nullptr);
        });
    }
    // Streaming (non-temporal) access to f and pi pays when an array of the
    // launch's members exceeds 96 MiB: the memory-side cache (256 MB) cannot
    // keep them until their next reader anyway.  Measured: 256^3 FP32 step
    // 0.938 -> 0.919 ms; at 128^3 (17 MB arrays) streaming cost 18 %.
    static bool streams(int members) { return (double)members * G::MS * sizeof(cpx<T>) > 96.0 * (1 << 20); }
    // Pass directions alternate so that each pass first reads what its
    // predecessor wrote last (still in the memory-side cache): z kick
    // forward, z inverse reversed, y inverse forward, x pass (reads every x
    // plane per block, no order to exploit), y forward reversed.  At 256^3
    // FP32 (134 MB arrays) this took the step from 1.007 to 0.939 ms.
    template <int SIGN>
    int launch_y(psc_ctx* c, cpx<T>* W, int members) {
        const double X = (double)members * G::N3 * sizeof(cpx<T>);
        static_assert(GY::N == N && std::is_same_v<typename GY::T, T>, "y plan of another lattice");
        dim3 grid(N / GY::TC, N, members), block(GY::THREADS_COL);
        return timed(c, SIGN > 0 ? K_Y_INV : K_Y_FWD, 2 * X, [&] {
            hipLaunchKernelGGL((k_strided<GY, 1, SIGN, false, (SIGN < 0)>), grid, block, 0, c->stream, W, P(c->tw));
        });
    }
    // Members m0 .. m0 + n - 1 (all of them by default).
    int launch_x_force(psc_ctx* c, int m0 = 0, int n = -1) {
        if (n < 0) n = c->B;
        const size_t off = (size_t)m0 * G::MS;
        const double X = (double)n * G::N3 * sizeof(cpx<T>);
        dim3 grid(N / G::TC, N, n), block(G::THREADS_COL);
        return timed(c, K_X_FORCE, 2 * X, [&] {
            hipLaunchKernelGGL((k_x_force<G>), grid, block, 0, c->stream, P(c->W) + off, c->ms + m0, P(c->tw),
                               c->partV + (size_t)m0 * G::BLOCKS_COL, P(c->Pf) + off);
        });
    }
    int launch_z_kick(psc_ctx* c, bool reduce, double c1, double c2, int use_a_start, int m0 = 0, int n = -1) {
        if (n < 0) n = c->B;
        const size_t off = (size_t)m0 * G::MS;
        double* pkg = c->partKG + (size_t)m0 * G::BLOCKS_ROW * 2;
        const double X = (double)n * G::N3 * sizeof(cpx<T>);
        dim3 grid(G::BLOCKS_ROW, n), block(G::THREADS_ROW);
        const StepConsts sc = consts_of(c);
        const bool st = streams(n);
        if (reduce)
            return timed(c, K_Z_FWD_KICK, 4 * X, [&] {
                if (st)
                    hipLaunchKernelGGL((k_z_kick<G, true, false, true>), grid, block, 0, c->stream, P(c->W) + off,
                                       P(c->Pf) + off, P(c->Pp) + off, c->ms + m0, P(c->tw), pkg, c1, c2,
                                       use_a_start, sc, nullptr);
                else
                    hipLaunchKernelGGL((k_z_kick<G, true>), grid, block, 0, c->stream, P(c->W) + off,
                                       P(c->Pf) + off, P(c->Pp) + off, c->ms + m0, P(c->tw), pkg, c1, c2,
                                       use_a_start, sc, nullptr);
            });
        return timed(c, K_Z_FWD_KICK_NR, 4 * X, [&] {
            if (st)
                hipLaunchKernelGGL((k_z_kick<G, false, false, true>), grid, block, 0, c->stream, P(c->W) + off,
                                   P(c->Pf) + off, P(c->Pp) + off, c->ms + m0, P(c->tw), pkg, c1, c2, use_a_start,
                                   sc, nullptr);
            else
                hipLaunchKernelGGL((k_z_kick<G, false>), grid, block, 0, c->stream, P(c->W) + off, P(c->Pf) + off,
                                   P(c->Pp) + off, c->ms + m0, P(c->tw), pkg, c1, c2, use_a_start, sc, nullptr);
        });
    }
    // Reduce the substep partials to the scalar update.  Passes with more than
    // kScalarDirect blocks per member are first cut to kPreChunks compensated
    // chunk sums by many blocks, so the single-block update stays short.
    int launch_scalars(psc_ctx* c, ScalarOp op, MemberState* rec = nullptr) {
        const StepConsts sc = consts_of(c);
        const double* pv = c->partV;
        const double* pkg = c->partKG;
        if (op.finish && (op.nV > kScalarDirect || op.nKG > kScalarDirect)) {
            hipLaunchKernelGGL(k_prereduce, dim3(2 * kPreChunks, c->B), dim3(256), 0, c->stream, c->partV, op.nV,
                               c->partV2, c->partKG, op.nKG, c->partKG2, kPreChunks);
            if (op.nV > kPreChunks) {
                pv = c->partV2;
                op.nV = kPreChunks;
            }
            if (op.nKG > kPreChunks) {
                pkg = c->partKG2;
                op.nKG = kPreChunks;
            }
        }
        const double bytes = (double)c->B * ((double)(op.nV + 2 * op.nKG) * 8.0 + 2 * sizeof(MemberState));
        return timed(c, K_SCALARS, bytes, [&] {
            hipLaunchKernelGGL(k_scalars, dim3(c->B), dim3(kScalarThreads), 0, c->stream, c->ms, pv, pkg, op, sc,
                               rec);
        });
    }

    // y inverse of the GW chain: fields and z gradient in place, y gradient
    // formed from the fields (read 2X, write 3X)
    int launch_y_inverse_gw(psc_ctx* c, size_t astride) {
        const double X = (double)c->B * G::N3 * sizeof(cpx<T>);
        return timed(c, K_Y_INV_GW, 5 * X, [&] {
            hipLaunchKernelGGL((k_y_inverse_gw<GY>), dim3(N / GY::TC, N, c->B), dim3(GY::THREADS_COL), 0, c->stream,
                               P(c->W), P(c->tw), astride, c->cfg.kIR);
        });
    }
    int force_chain_gw(psc_ctx* c, bool drift, bool reduce, double c1, double c2, int use_a_start) {
        const size_t astride = (size_t)c->B * G::MS;
        const double X = (double)c->B * G::N3 * sizeof(cpx<T>);
        const StepConsts sc = consts_of(c);
        int rc;
        dim3 rgrid(G::BLOCKS_ROW, c->B), rblock(G::THREADS_ROW);
        if (drift)
            rc = timed(c, K_Z_INV_GW, 14 * X, [&] {
                hipLaunchKernelGGL((k_z_inverse_gw<G, true>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pp),
                                   P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), astride, c->cfg.kIR);
            });
        else
            rc = timed(c, K_Z_INV_GW, 3 * X, [&] {
                hipLaunchKernelGGL((k_z_inverse_gw<G, false>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pp),
                                   P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), astride, c->cfg.kIR);
            });
        if (rc) return rc;
        if ((rc = launch_y_inverse_gw(c, astride))) return rc;
        // gradients: x inverse, the source and its x forward transforms in
        // one pass, before the fields' x pass overwrites W0 with V'
        rc = timed(c, K_X_FORCE_GW, 6 * X, [&] {
            hipLaunchKernelGGL((k_x_source<G>), dim3(N / G::TC, N, c->B), dim3(G::THREADS_COL), 0, c->stream,
                               P(c->W), P(c->tw), astride, c->cfg.kIR, nullptr);
        });
        if (rc) return rc;
        if ((rc = launch_x_force(c))) return rc;
        if ((rc = launch_y<-1>(c, P(c->W), 4 * c->B))) return rc;
        if (reduce)
            return timed(c, K_Z_KICK_GW, 16 * X, [&] {
                hipLaunchKernelGGL((k_z_kick_gw<G, true>), rgrid, rblock, 0, c->stream, P(c->W), P(c->Pf),
                                   P(c->Pp), P(c->U), P(c->PU), c->ms, P(c->tw), c->partKG, c1, c2, use_a_start,
                                   sc, astride);
            });
        return timed(c, K_Z_KICK_GW, 16 * X, [&] {
            hipLaunchKernelGGL((k_z_kick_gw<G, false>), rgrid, rblock, 0, c->stream, P(c->W), P(c->Pf),
                               P(c->Pp), P(c->U), P(c->PU), c->ms, P(c->tw), c->partKG, c1, c2, use_a_start, sc,
                               astride);
        });
    }

    // One force evaluation chain; DRIFT applies the substep's drift first.
    int force_chain(psc_ctx* c, bool drift, bool reduce, double c1, double c2, int use_a_start) {
        if (c->padeng) return c->padeng->force_chain(c, drift, reduce, c1, c2, use_a_start);
        if (c->cfg.with_gw) return force_chain_gw(c, drift, reduce, c1, c2, use_a_start);
        // An ensemble runs the chain on groups of members whose arrays fit the
        // memory-side cache together, so that each pass of a group reads what
        // the group's previous pass has just written.
        const int gs = c->member_group > 0 ? std::min(c->member_group, c->B) : c->B;
        for (int m0 = 0; m0 < c->B; m0 += gs) {
            const int n = std::min(gs, c->B - m0);
            const size_t off = (size_t)m0 * G::MS;
            int rc;
            if ((rc = launch_z_inverse(c, drift, P(c->Pf) + off, P(c->Pf) + off, P(c->Pp) + off, P(c->W) + off,
                                       c->ms + m0, n)))
                return rc;
            if ((rc = launch_y<+1>(c, P(c->W) + off, n))) return rc;
            if ((rc = launch_x_force(c, m0, n))) return rc;
            if ((rc = launch_y<-1>(c, P(c->W) + off, n))) return rc;
            if ((rc = launch_z_kick(c, reduce, c1, c2, use_a_start, m0, n))) return rc;
        }
        return PSC_OK;
    }

    // All substeps of one step; `final` drops the trailing first-kick of the
    // next substep so the momenta are synchronised on return.
    int record_step(psc_ctx* c, bool final) {
        const int S = (int)c->w.size();
        const double dt = c->cfg.dt;
        for (int p = 0; p < S; ++p) {
            const bool last = final && p == S - 1;
            const double wn = c->w[(p + 1) % S];
            int rc = force_chain(c, true, true, c->w[p] * dt / 2.0, last ? 0.0 : wn * dt / 2.0, 0);
            if (rc) return rc;
            ScalarOp op{};
            op.finish = 1;
            op.last_of_step = p == S - 1;
            op.begin = last ? 0 : 1;
            op.w_finish = c->w[p];
            op.w_begin = wn;
            op.nV = c->padeng ? c->padeng->nV() : nV_blocks(c);
            op.nKG = c->padeng ? c->padeng->nKG() : G::BLOCKS_ROW;
            if ((rc = launch_scalars(c, op))) return rc;
        }
        return PSC_OK;
    }

    int build_graph(psc_ctx* c, bool final, hipGraphExec_t* out) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = record_step(c, final);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return c->fail(PSC_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        HIPCHK(c, hipGraphInstantiate(out, g, nullptr, nullptr, 0));
        HIPCHK(c, hipGraphDestroy(g));
        return PSC_OK;
    }

    int step(psc_ctx* c, int nsteps) override {
        if (nsteps <= 0) return PSC_OK;
        ScalarOp op{};
        op.begin = 1;
        op.w_begin = c->w[0];
        op.nV = c->padeng ? c->padeng->nV() : nV_blocks(c);
        op.nKG = c->padeng ? c->padeng->nKG() : G::BLOCKS_ROW;
        int rc = launch_scalars(c, op);
        if (rc) return rc;
        if ((rc = force_chain(c, false, false, c->w[0] * c->cfg.dt / 2.0, 0.0, 1))) return rc;
        const bool graph = c->cfg.use_graph && !c->timing;
        if (graph) {
            if (!c->g_mid && (rc = build_graph(c, false, &c->g_mid))) return rc;
            if (!c->g_last && (rc = build_graph(c, true, &c->g_last))) return rc;
            for (int s = 0; s < nsteps - 1; ++s) HIPCHK(c, hipGraphLaunch(c->g_mid, c->stream));
            HIPCHK(c, hipGraphLaunch(c->g_last, c->stream));
        } else {
            for (int s = 0; s < nsteps; ++s)
                if ((rc = record_step(c, s == nsteps - 1))) return rc;
        }
        HIPCHK(c, hipGetLastError());   // a failed launch is reported here, not by a later call
        return c->timing ? drain_timing(c) : PSC_OK;
    }

    int real_space(psc_ctx* c, int member, const void* src) override {
        const size_t off = (size_t)member * G::MS;
        cpx<T>* W = P(c->W) + off;
        int rc = launch_z_inverse(c, false, static_cast<const cpx<T>*>(src) + off, nullptr, nullptr, W,
                                  c->ms + member, 1);
        if (rc) return rc;
        if ((rc = launch_y<+1>(c, W, 1))) return rc;
        dim3 grid(N / G::TC, N, 1), block(G::THREADS_COL);
        rc = timed(c, K_DIAG, 2.0 * G::N3 * sizeof(cpx<T>), [&] {
            hipLaunchKernelGGL((k_x_inverse_diag<G>), grid, block, 0, c->stream, W, c->ms + member,
                               P(c->tw), c->part, static_cast<const cpx<T>*>(src) + off);
        });
        if (rc) return rc;
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, c->stream, c->part, G::BLOCKS_COL, 6,
                           c->sums);
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    int field_sums(psc_ctx* c, int member) override {
        const size_t off = (size_t)member * G::MS;
        const int nb = 256;
        hipLaunchKernelGGL((k_field_sums<T, N>), dim3(nb, 1), dim3(256), 0, c->stream, P(c->Pf) + off,
                           P(c->Pp) + off, c->part, nb);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, c->stream, c->part, nb, 4, c->sums);
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    int spectra(psc_ctx* c, int member, int fld) override {
        const size_t off = (size_t)member * G::MS;
        const int nbt = (int)(std::sqrt(3.0) * (N / 2)) + 2;
        const int nb = 128;
        const size_t shmem = (size_t)nbt * 4 * sizeof(double);
        hipLaunchKernelGGL((k_spectra<T, N>), dim3(nb, 1), dim3(256), shmem, c->stream, P(c->Pf) + off,
                           P(c->Pp) + off, c->ms + member, fld, c->cfg.alpha, nbt, c->part, nb);
        const int width = nbt * 4;
        hipLaunchKernelGGL(k_sum_partials, dim3((width + 255) / 256), dim3(256), 0, c->stream, c->part,
                           nb, width, c->sums);
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    int gw_spectra(psc_ctx* c, int member, int projection, int nlog) override;
    int adjoint(psc_ctx* c, const psc_adjoint_spec* sp, double* out, double* dchi) override;

    int to_spectrum(psc_ctx* c, int member, void* dst) override {
        const size_t off = (size_t)member * G::MS;
        cpx<T>* W = P(c->W) + off;
        dim3 cgrid(N / G::TC, N, 1), cblock(G::THREADS_COL);
        hipLaunchKernelGGL((k_strided<G, 0, -1>), cgrid, cblock, 0, c->stream, W, P(c->tw));
        int rc = launch_y<-1>(c, W, 1);
        if (rc) return rc;
        hipLaunchKernelGGL((k_z_forward<G>), dim3(G::BLOCKS_ROW, 1), dim3(G::THREADS_ROW), 0, c->stream, W,
                           static_cast<cpx<T>*>(dst) + off, P(c->tw));
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    int density(psc_ctx* c, int member, int kind, double* D, double scale) override {
        const size_t off = (size_t)member * G::MS;
        cpx<T>* W = P(c->W) + off;
        const cpx<T>* tw = P(c->tw);
        dim3 cgrid(N / G::TC, N, 1), cblock(G::THREADS_COL);
        int rc;
        if (kind < 2) {
            const cpx<T>* src = P(kind == 0 ? c->Pf : c->Pp) + off;
            if ((rc = launch_z_inverse(c, false, src, nullptr, nullptr, W, c->ms + member, 1))) return rc;
            if ((rc = launch_y<+1>(c, W, 1))) return rc;
            if (kind == 0)
                hipLaunchKernelGGL((k_x_density<G, 0>), cgrid, cblock, 0, c->stream, W, D, c->ms + member, tw,
                                   src, scale);
            else
                hipLaunchKernelGGL((k_x_density<G, 1>), cgrid, cblock, 0, c->stream, W, D, c->ms + member, tw,
                                   src, scale);
        } else {
            HIPCHK(c, hipMemsetAsync(D, 0, G::N3 * sizeof(double), c->stream));
            for (int d = 0; d < 3; ++d) {
                hipLaunchKernelGGL((k_z_inverse_deriv<G>), dim3(G::BLOCKS_ROW, 1), dim3(G::THREADS_ROW), 0,
                                   c->stream, P(c->Pf) + off, W, tw, d, c->cfg.kIR);
                if ((rc = launch_y<+1>(c, W, 1))) return rc;
                hipLaunchKernelGGL((k_x_density<G, 2>), cgrid, cblock, 0, c->stream, W, D, c->ms + member, tw,
                                   nullptr, scale);
            }
        }
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    // ---- adjoint helpers --------------------------------------------------
    static constexpr int kObjBlocks = (int)(((size_t)N * N * (N / 2 + 1) + 255) / 256);
    int adj_workspace(psc_ctx* c, int K, int nseg, bool gw, AdjWork** out) {
        const int S = (int)c->w.size();
        AdjWork* cur = c->adj.get();
        if (cur && cur->K == K && cur->S == S && cur->nseg >= nseg && cur->gw == gw) {
            *out = cur;
            return PSC_OK;
        }
        c->adj.reset();
        auto w = std::make_unique<AdjWork>();
        const size_t arr = G::MS * sizeof(cpx<T>);
        const int SK = K * S;
        bool ok = true;
        auto dalloc = [&](size_t bytes) -> void* {
            void* q = nullptr;
            if (!ok || hipMalloc(&q, bytes) != hipSuccess) {
                ok = false;
                return nullptr;
            }
            w->owned.push_back(q);
            w->bytes += (double)bytes;
            return q;
        };
        w->K = K;
        w->S = S;
        w->nseg = nseg;
        w->ck.resize(2 * (size_t)nseg);
        for (auto& q : w->ck) q = dalloc(arr);
        w->ck_ms = static_cast<MemberState*>(dalloc(sizeof(MemberState) * nseg));
        w->f.resize(SK + 1);
        w->p.resize(SK + 1);
        w->r.resize(SK + 1);
        w->ph.resize(SK);
        for (int j = 0; j <= SK; ++j) {
            w->f[j] = dalloc(arr);
            w->p[j] = dalloc(arr);
            w->r[j] = dalloc(arr);
            if (j < SK) w->ph[j] = dalloc(arr);
        }
        w->rms = static_cast<MemberState*>(dalloc(sizeof(MemberState) * (SK + 1)));
        w->fbar = dalloc(arr);
        w->pbar = dalloc(arr);
        w->st = static_cast<AdjState*>(dalloc(sizeof(AdjState)));
        w->co = static_cast<AdjCoef*>(dalloc(sizeof(AdjCoef)));
        w->p1 = static_cast<double*>(dalloc(sizeof(double) * G::BLOCKS_ROW));
        w->p3 = static_cast<double*>(dalloc(sizeof(double) * 3 * G::BLOCKS_COL));
        w->p5 = static_cast<double*>(dalloc(sizeof(double) * G::BLOCKS_ROW));
        w->gw = gw;
        if (gw) {
            w->ck_u.resize(nseg);
            w->ck_pu.resize(nseg);
            for (int sg = 0; sg < nseg; ++sg) {
                w->ck_u[sg] = dalloc(3 * arr);
                w->ck_pu[sg] = dalloc(3 * arr);
            }
            w->u.resize(SK + 1);
            w->g.resize(SK + 1);
            w->puh.resize(SK);
            for (int j = 0; j <= SK; ++j) {
                w->u[j] = dalloc(3 * arr);
                w->g[j] = dalloc(3 * arr);
                if (j < SK) w->puh[j] = dalloc(3 * arr);
            }
            w->ubar = dalloc(3 * arr);
            w->pubar = dalloc(3 * arr);
            w->p1g = static_cast<double*>(dalloc(sizeof(double) * G::BLOCKS_ROW));
            w->p3g = static_cast<double*>(dalloc(sizeof(double) * G::BLOCKS_COL));
            w->p5g = static_cast<double*>(dalloc(sizeof(double) * G::BLOCKS_ROW));
            w->pobj = static_cast<double*>(dalloc(sizeof(double) * (kObjBlocks + 2)));
        }
        if (!ok) return c->fail(PSC_ERR_NOMEM, "adjoint: device allocation failed");
        c->adj = std::move(w);
        *out = c->adj.get();
        return PSC_OK;
    }

    // Replays `steps` steps from the current state, recording f, pi and the
    // real-space fields at every substep boundary, pi_half per substep and the
    // scalar state after each scalar update.
    int record_segment(psc_ctx* c, int steps, AdjWork& w) {
        if (w.gw) return record_segment_gw(c, steps, w);
        const int S = (int)c->w.size();
        const double dt = c->cfg.dt;
        const StepConsts sc = consts_of(c);
        HIPCHK(c, hipMemcpyAsync(w.p[0], c->Pp, G::MS * sizeof(cpx<T>), hipMemcpyDeviceToDevice, c->stream));
        ScalarOp op{};
        op.begin = 1;
        op.w_begin = c->w[0];
        op.nV = G::BLOCKS_COL;
        op.nKG = G::BLOCKS_ROW;
        int rc = launch_scalars(c, op, w.rms);
        if (rc) return rc;
        dim3 rgrid(G::BLOCKS_ROW, 1), rblock(G::THREADS_ROW), cgrid(N / G::TC, N, 1), cblock(G::THREADS_COL);
        // prologue force at (f_0, a_0): first kick, real-space fields of f_0
        hipLaunchKernelGGL((k_z_inverse<G, false, false, true>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pf),
                           P(c->Pp), P(c->W), c->ms, P(c->tw), nullptr, nullptr);
        if ((rc = launch_y<+1>(c, P(c->W), 1))) return rc;
        hipLaunchKernelGGL((k_x_force<G, true>), cgrid, cblock, 0, c->stream, P(c->W), c->ms, P(c->tw), c->partV,
                           P(c->Pf), P(w.r[0]));
        if ((rc = launch_y<-1>(c, P(c->W), 1))) return rc;
        hipLaunchKernelGGL((k_z_kick<G, false>), rgrid, rblock, 0, c->stream, P(c->W), P(c->Pf), P(c->Pp), c->ms,
                           P(c->tw), c->partKG, c->w[0] * dt / 2.0, 0.0, 1, sc, nullptr);
        for (int j = 0; j < steps * S; ++j) {
            const int p = j % S;
            const bool last = j == steps * S - 1;
            const double wn = c->w[(p + 1) % S];
            hipLaunchKernelGGL((k_z_inverse<G, true, true, true>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pf),
                               P(c->Pp), P(c->W), c->ms, P(c->tw), P(w.f[j]), P(w.ph[j]));
            if ((rc = launch_y<+1>(c, P(c->W), 1))) return rc;
            hipLaunchKernelGGL((k_x_force<G, true>), cgrid, cblock, 0, c->stream, P(c->W), c->ms, P(c->tw),
                               c->partV, P(c->Pf), P(w.r[j + 1]));
            if ((rc = launch_y<-1>(c, P(c->W), 1))) return rc;
            hipLaunchKernelGGL((k_z_kick<G, true, true>), rgrid, rblock, 0, c->stream, P(c->W), P(c->Pf),
                               P(c->Pp), c->ms, P(c->tw), c->partKG, c->w[p] * dt / 2.0, last ? 0.0 : wn * dt / 2.0,
                               0, sc, P(w.p[j + 1]));
            ScalarOp o{};
            o.finish = 1;
            o.last_of_step = p == S - 1;
            o.begin = last ? 0 : 1;
            o.w_finish = c->w[p];
            o.w_begin = wn;
            o.nV = G::BLOCKS_COL;
            o.nKG = G::BLOCKS_ROW;
            if ((rc = launch_scalars(c, o, w.rms + j + 1))) return rc;
        }
        HIPCHK(c, hipMemcpyAsync(w.f[steps * S], c->Pf, G::MS * sizeof(cpx<T>), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    // The same replay through the GW chain (force_chain_gw), recording also u
    // and the real-space gradients at every boundary and pi_u half per substep.
    int record_segment_gw(psc_ctx* c, int steps, AdjWork& w) {
        const int S = (int)c->w.size();
        const double dt = c->cfg.dt;
        const StepConsts sc = consts_of(c);
        const size_t arr = G::MS * sizeof(cpx<T>), as = G::MS;   // one member: arrays MS apart
        const double kIR = c->cfg.kIR;
        HIPCHK(c, hipMemcpyAsync(w.p[0], c->Pp, arr, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(w.u[0], c->U, 3 * arr, hipMemcpyDeviceToDevice, c->stream));
        ScalarOp op{};
        op.begin = 1;
        op.w_begin = c->w[0];
        op.nV = G::BLOCKS_COL;
        op.nKG = G::BLOCKS_ROW;
        int rc = launch_scalars(c, op, w.rms);
        if (rc) return rc;
        dim3 rgrid(G::BLOCKS_ROW, 1), rblock(G::THREADS_ROW), cgrid(N / G::TC, N, 1), cblock(G::THREADS_COL);
        auto middle = [&](int jb) -> int {   // y inverse .. y forward of the GW chain at boundary jb
            int r2 = launch_y_inverse_gw(c, as);
            if (r2) return r2;
            hipLaunchKernelGGL((k_x_source<G, true>), cgrid, cblock, 0, c->stream, P(c->W), P(c->tw), as, kIR,
                               P(w.g[jb]));
            hipLaunchKernelGGL((k_x_force<G, true>), cgrid, cblock, 0, c->stream, P(c->W), c->ms, P(c->tw),
                               c->partV, P(c->Pf), P(w.r[jb]));
            return launch_y<-1>(c, P(c->W), 4);
        };
        // prologue force at (f_0, a_0, u_0): first kicks
        hipLaunchKernelGGL((k_z_inverse_gw<G, false>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pp), P(c->W),
                           P(c->U), P(c->PU), c->ms, P(c->tw), as, kIR, nullptr, nullptr, nullptr, nullptr);
        if ((rc = middle(0))) return rc;
        hipLaunchKernelGGL((k_z_kick_gw<G, false>), rgrid, rblock, 0, c->stream, P(c->W), P(c->Pf), P(c->Pp),
                           P(c->U), P(c->PU), c->ms, P(c->tw), c->partKG, c->w[0] * dt / 2.0, 0.0, 1, sc, as,
                           nullptr);
        for (int j = 0; j < steps * S; ++j) {
            const int p = j % S;
            const bool last = j == steps * S - 1;
            const double wn = c->w[(p + 1) % S];
            hipLaunchKernelGGL((k_z_inverse_gw<G, true, true>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pp),
                               P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), as, kIR, P(w.f[j]), P(w.ph[j]),
                               P(w.u[j + 1]), P(w.puh[j]));
            if ((rc = middle(j + 1))) return rc;
            hipLaunchKernelGGL((k_z_kick_gw<G, true, true>), rgrid, rblock, 0, c->stream, P(c->W), P(c->Pf),
                               P(c->Pp), P(c->U), P(c->PU), c->ms, P(c->tw), c->partKG, c->w[p] * dt / 2.0,
                               last ? 0.0 : wn * dt / 2.0, 0, sc, as, P(w.p[j + 1]));
            ScalarOp o{};
            o.finish = 1;
            o.last_of_step = p == S - 1;
            o.begin = last ? 0 : 1;
            o.w_finish = c->w[p];
            o.w_begin = wn;
            o.nV = G::BLOCKS_COL;
            o.nKG = G::BLOCKS_ROW;
            if ((rc = launch_scalars(c, o, w.rms + j + 1))) return rc;
        }
        HIPCHK(c, hipMemcpyAsync(w.f[steps * S], c->Pf, arr, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    // One backward substep (or the t = 0 pass): P1, y inverse, P3, y forward,
    // P5; with the GW objective P1g and P3g run beside P1 and P3, the y
    // forward also merges the gradient adjoint's x and y shares into W0, and
    // P5 adds its z share to fbar before the drift adjoint.
    int backward_passes(psc_ctx* c, AdjWork& w, int jf, int jph, bool gw) {
        dim3 rgrid(G::BLOCKS_ROW, 1), rblock(G::THREADS_ROW), cgrid(N / G::TC, N, 1), cblock(G::THREADS_COL);
        const size_t as = G::MS;
        hipLaunchKernelGGL((k_adj_p1<G>), rgrid, rblock, 0, c->stream, P(w.f[jf]), P(w.p[jf]), P(w.fbar),
                           P(w.pbar), P(c->W), w.co, P(c->tw), w.p1);
        if (gw)
            hipLaunchKernelGGL((k_adj_p1g<G>), rgrid, rblock, 0, c->stream, P(w.pubar), P(w.ubar), P(w.u[jf]),
                               P(c->W), as, w.co, P(c->tw), w.p1g);
        int rc = launch_y<+1>(c, P(c->W), gw ? 4 : 1);
        if (rc) return rc;
        hipLaunchKernelGGL((k_adj_p3<G>), cgrid, cblock, 0, c->stream, P(c->W), P(w.r[jf]), c->ms, w.co, P(c->tw),
                           w.p3);
        if (gw)
            hipLaunchKernelGGL((k_adj_p3g<G>), cgrid, cblock, 0, c->stream, P(c->W) + as, P(c->W), P(w.g[jf]), as,
                               c->cfg.kIR, P(c->tw), w.p3g);
        if (gw)
            hipLaunchKernelGGL((k_adj_y_fwd_gw<GY>), dim3(N / GY::TC, N, 1), dim3(GY::THREADS_COL), 0, c->stream,
                               P(c->W), as, c->cfg.kIR, P(c->tw));
        else if ((rc = launch_y<-1>(c, P(c->W), 1)))
            return rc;
        if (gw)
            hipLaunchKernelGGL((k_adj_p5<G, true>), rgrid, rblock, 0, c->stream, P(c->W), P(w.fbar), P(w.pbar),
                               P(w.ph[jph]), w.co, P(c->tw), w.p5, as, c->cfg.kIR, P(w.ubar), P(w.pubar),
                               P(w.puh[jph]), w.p5g);
        else
            hipLaunchKernelGGL((k_adj_p5<G>), rgrid, rblock, 0, c->stream, P(c->W), P(w.fbar), P(w.pbar),
                               P(w.ph[jph]), w.co, P(c->tw), w.p5);
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    AdjPartials partials(const AdjWork& w) const { return AdjPartials{w.p1, w.p3, w.p5, w.p1g, w.p3g, w.p5g}; }

    AdjScalarOp adj_op(const psc_ctx* c) {
        const StepConsts sc = consts_of(c);
        AdjScalarOp op{};
        op.static_bg = sc.expansion_static;
        op.alpha = sc.alpha;
        op.kIR = sc.kIR;
        op.f2m = sc.f2m;
        op.N3 = sc.N3;
        op.n1 = G::BLOCKS_ROW;
        op.n3 = G::BLOCKS_COL;
        op.n5 = G::BLOCKS_ROW;
        op.gw = c->adj && c->adj->gw ? 1 : 0;
        return op;
    }

    // Sweeps the recorded segment of `steps` steps backwards.
    int backward_segment(psc_ctx* c, int steps, AdjWork& w) {
        const int S = (int)c->w.size();
        const int nsub = steps * S;
        AdjScalarOp op = adj_op(c);
        for (int j = nsub - 1; j >= 0; --j) {
            op.mode = 1 | 2;
            op.h_pre = c->w[j % S] * c->cfg.dt / 2.0;
            op.end_of_segment = j == nsub - 1;
            hipLaunchKernelGGL(k_adj_scalars, dim3(1), dim3(256), 0, c->stream, w.st, w.co, partials(w),
                               w.rms + j + 1, w.rms + j, op);
            int rc = backward_passes(c, w, j + 1, j, w.gw);
            if (rc) return rc;
        }
        return PSC_OK;
    }

    // Records (bwd = false) or sweeps (bwd = true) a segment of `steps` steps;
    // full segments go through a graph captured on first use when steps are
    // graph-captured.
    int run_segment(psc_ctx* c, int steps, AdjWork& w, bool bwd) {
        if (steps != w.K || !c->cfg.use_graph)
            return bwd ? backward_segment(c, steps, w) : record_segment(c, steps, w);
        hipGraphExec_t& ge = bwd ? w.g_bwd : w.g_rec;
        if (!ge) {
            hipGraph_t g;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = bwd ? backward_segment(c, steps, w) : record_segment(c, steps, w);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess)
                return c->fail(PSC_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            HIPCHK(c, hipGraphDestroy(g));
        }
        HIPCHK(c, hipGraphLaunch(ge, c->stream));
        return PSC_OK;
    }
};


template <typename T, int N>
int gw_spectra_impl(psc_ctx* c, int member, int projection, int nlog) {
    const size_t off = (size_t)member * member_stride(N);
    const int nbt = (int)(std::sqrt(3.0) * (N / 2)) + 2;
    const int nb = 128;
    const int width = 3 * nbt + nlog;
    const size_t shmem = (size_t)width * sizeof(double);
    hipLaunchKernelGGL((k_gw_spectra<T, N>), dim3(nb), dim3(256), shmem, c->stream,
                       static_cast<cpx<T>*>(c->U) + off, static_cast<cpx<T>*>(c->PU) + off,
                       (size_t)c->B * member_stride(N), projection, nbt, nlog, c->part, nb);
    hipLaunchKernelGGL(k_sum_partials, dim3((width + 255) / 256), dim3(256), 0, c->stream, c->part, nb, width,
                       c->sums);
    HIPCHK(c, hipGetLastError());
    return PSC_OK;
}

template <class G, class GY>
int Engine<G, GY>::gw_spectra(psc_ctx* c, int member, int projection, int nlog) {
    return gw_spectra_impl<typename G::T, G::N>(c, member, projection, nlog);
}

// The dealiased stepper's passes on the padded shape GP (N_pad = GP::N = 2N);
// the two z-axis row passes (padding, extraction) may run on another plan GR.
template <class GP, class GR = GP>
struct PadEngine final : PadBase {
    using T = typename GP::T;
    static constexpr int NP = GP::N, N = NP / 2;
    static_assert(GR::N == NP && std::is_same_v<typename GR::T, T>, "row plan of another lattice");
    static cpx<T>* P(void* p) { return static_cast<cpx<T>*>(p); }
    static constexpr double XN = (double)N * N * N * sizeof(cpx<T>);
    static constexpr double XP = 8.0 * XN;
    static constexpr double KP = (N + 1.0) / (2.0 * N);   // share of a padded axis holding modes

    int nV() const override { return GP::BLOCKS_COL; }
    int nKG() const override { return (int)((size_t)N * N * N / 2048); }
    size_t padded_elems() const override { return GP::MS; }

    // padding with the z-axis inverse, then the y-axis inverse, then V' and
    // the sum of V in the x pass
    int front(psc_ctx* c) {
        int rc = timed(c, K_PAD_Z_INV, XN + XP * KP * KP, [&] {
            hipLaunchKernelGGL((k_pad_z<GR>), dim3(GR::BLOCKS_ROW), dim3(GR::THREADS_ROW), 0, c->stream, P(c->Pf),
                               P(c->Gf), P(c->Wp), P(c->dcp), P(c->twp));
        });
        if (rc) return rc;
        // the y passes and the stored x rows cover the padded x planes |x| <= N/2;
        // the y inverse reads the rows |y| <= N/2 of them
        rc = timed(c, K_PAD_Y_INV, XP * KP * (KP + 1.0), [&] {
            hipLaunchKernelGGL((k_strided<GP, 1, +1, true>), dim3(NP / GP::TC, N + 1, 1), dim3(GP::THREADS_COL), 0,
                               c->stream, P(c->Wp), P(c->twp));
        });
        if (rc) return rc;
        return timed(c, K_PAD_X_FORCE, XP * KP * 2.0, [&] {
            hipLaunchKernelGGL((k_x_force<GP, false, true>), dim3(NP / GP::TC, NP, 1), dim3(GP::THREADS_COL), 0,
                               c->stream, P(c->Wp), c->ms, P(c->twp), c->partV, P(c->dcp));
        });
    }

    int force_chain(psc_ctx* c, bool drift, bool reduce, double c1, double c2, int use_a_start) override {
        int rc;
        if (drift && (rc = timed(c, K_PAD_DRIFT, 3 * XN, [&] {
                hipLaunchKernelGGL((k_drift_pad<T, N>), dim3((unsigned)((size_t)N * N * N / 256)), dim3(256), 0,
                                   c->stream, P(c->Pf), P(c->Pp), P(c->Gf), P(c->Gp), c->ms);
            })))
            return rc;
        if ((rc = front(c))) return rc;
        rc = timed(c, K_PAD_Y_FWD, XP * KP * (1.0 + KP), [&] {
            hipLaunchKernelGGL((k_strided<GP, 1, -1, true>), dim3(NP / GP::TC, N + 1, 1), dim3(GP::THREADS_COL), 0,
                               c->stream, P(c->Wp), P(c->twp));
        });
        if (rc) return rc;
        rc = timed(c, K_PAD_Z_FWD, XN * (N + 1) * (N + 1) / ((double)N * N) * 3, [&] {
            hipLaunchKernelGGL((k_unpad_z<GR>), dim3(GR::BLOCKS_ROW), dim3(GR::THREADS_ROW), 0, c->stream, P(c->Wp),
                               P(c->Rp), P(c->twp));
        });
        if (rc) return rc;
        const StepConsts sc = consts_of(c);
        return timed(c, K_PAD_KICK, 4 * XN, [&] {
            if (reduce)
                hipLaunchKernelGGL((k_kick_pad<GP, true>), dim3(nKG()), dim3(256), 0, c->stream, P(c->Rp), P(c->Pf),
                                   P(c->Pp), P(c->Gf), P(c->Gp), c->ms, c->partKG, c1, c2, use_a_start, sc);
            else
                hipLaunchKernelGGL((k_kick_pad<GP, false>), dim3(nKG()), dim3(256), 0, c->stream, P(c->Rp), P(c->Pf),
                                   P(c->Pp), P(c->Gf), P(c->Gp), c->ms, c->partKG, c1, c2, use_a_start, sc);
        });
    }

    int mean_V(psc_ctx* c, double* V) override {
        int rc = front(c);
        if (rc) return rc;
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, c->stream, c->partV, nV(), 1, c->sums);
        HIPCHK(c, hipGetLastError());
        double v = 0.0;
        HIPCHK(c, hipMemcpyAsync(&v, c->sums, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *V = v / (8.0 * N * N * N);
        return PSC_OK;
    }
};

// Launch shapes: index 0 is the default of each (precision, N); further
// entries are the candidates the autotuner times (same results up to the
// rounding of a different radix plan; identical when only TC / L differ).
template <typename T, int N, class R, int TC, int L>
using S = Shape<T, N, R, TC, L>;

struct ShapeEntry {
    int precision, N;
    const char* name;
    EngineBase* (*make)();    // the stepper of the N lattice, or null (padded stepper only)
    PadBase* (*make_pad)();   // the dealiased stepper of the N/2 lattice, or null
};

template <class G, class GY = G>
EngineBase* make_one() { return new Engine<G, GY>(); }

template <class G, class GR = G>
PadBase* make_pad_one() {
    if constexpr (G::N >= 32) return new PadEngine<G, GR>();
    else return nullptr;
}

#define PSC_SHAPE(P, T, N, NAME, ...) {P, N, NAME, &make_one<S<T, N, __VA_ARGS__>>, &make_pad_one<S<T, N, __VA_ARGS__>>}
// A shape whose y passes run on the second plan GY; not offered to the
// dealiased stepper, whose y passes are its own.
#define PSC_SHAPE_Y(P, NAME, G, GY) {P, G::N, NAME, &make_one<G, GY>, nullptr}
// A shape of the padded stepper only, its z row passes on GR.
#define PSC_SHAPE_PAD(P, NAME, G, GR) {P, G::N, NAME, nullptr, &make_pad_one<G, GR>}
using F512_888 = S<float, 512, Radices<8, 8, 8>, 16, 4>;
using F512_Y = S<float, 512, Radices<16, 16, 2>, 32, 4>;
using F512_1616 = S<float, 512, Radices<16, 16, 2>, 16, 8>;
using D512_888 = S<double, 512, Radices<8, 8, 8>, 8, 4>;
using D512_Y = S<double, 512, Radices<16, 16, 2>, 16, 4>;
const ShapeEntry kShapes[] = {
    PSC_SHAPE(PSC_FP32, float, 16, "r4.4_tc16_l64", Radices<4, 4>, 16, 64),
    PSC_SHAPE(PSC_FP32, float, 32, "r8.4_tc32_l64", Radices<8, 4>, 32, 64),
    PSC_SHAPE(PSC_FP32, float, 64, "r8.8_tc32_l32", Radices<8, 8>, 32, 32),
    PSC_SHAPE(PSC_FP32, float, 128, "r16.8_tc32_l32", Radices<16, 8>, 32, 32),
    PSC_SHAPE(PSC_FP32, float, 128, "r16.8_tc16_l16", Radices<16, 8>, 16, 16),
    PSC_SHAPE(PSC_FP32, float, 128, "r8.4.4_tc32_l16", Radices<8, 4, 4>, 32, 16),
    PSC_SHAPE(PSC_FP32, float, 256, "r16.16_tc32_l16", Radices<16, 16>, 32, 16),
    PSC_SHAPE(PSC_FP32, float, 256, "r16.16_tc16_l16", Radices<16, 16>, 16, 16),
    PSC_SHAPE(PSC_FP32, float, 256, "r16.16_tc32_l8", Radices<16, 16>, 32, 8),
    PSC_SHAPE(PSC_FP32, float, 256, "r8.8.4_tc16_l8", Radices<8, 8, 4>, 16, 8),
    PSC_SHAPE_PAD(PSC_FP32, "r16.16.2_tc16_l8_z8.8.8_l4", F512_1616, F512_888),
    PSC_SHAPE_Y(PSC_FP32, "r8.8.8_tc16_l4_y16.16.2_tc32", F512_888, F512_Y),
    PSC_SHAPE(PSC_FP32, float, 512, "r8.8.8_tc16_l4", Radices<8, 8, 8>, 16, 4),
    PSC_SHAPE(PSC_FP32, float, 512, "r16.16.2_tc16_l8", Radices<16, 16, 2>, 16, 8),
    PSC_SHAPE(PSC_FP32, float, 512, "r8.8.8_tc8_l8", Radices<8, 8, 8>, 8, 8),
    PSC_SHAPE(PSC_FP64, double, 16, "r4.4_tc16_l64", Radices<4, 4>, 16, 64),
    PSC_SHAPE(PSC_FP64, double, 32, "r8.4_tc32_l64", Radices<8, 4>, 32, 64),
    PSC_SHAPE(PSC_FP64, double, 64, "r8.8_tc32_l32", Radices<8, 8>, 32, 32),
    PSC_SHAPE(PSC_FP64, double, 128, "r16.8_tc16_l16", Radices<16, 8>, 16, 16),
    PSC_SHAPE(PSC_FP64, double, 128, "r8.4.4_tc16_l16", Radices<8, 4, 4>, 16, 16),
    PSC_SHAPE(PSC_FP64, double, 256, "r16.16_tc16_l8", Radices<16, 16>, 16, 8),
    PSC_SHAPE(PSC_FP64, double, 256, "r8.8.4_tc16_l8", Radices<8, 8, 4>, 16, 8),
    PSC_SHAPE(PSC_FP64, double, 256, "r16.16_tc8_l8", Radices<16, 16>, 8, 8),
    PSC_SHAPE_Y(PSC_FP64, "r8.8.8_tc8_l4_y16.16.2_tc16", D512_888, D512_Y),
    PSC_SHAPE(PSC_FP64, double, 512, "r8.8.8_tc8_l4", Radices<8, 8, 8>, 8, 4),
    PSC_SHAPE(PSC_FP64, double, 512, "r16.16.2_tc8_l4", Radices<16, 16, 2>, 8, 4),
};
#undef PSC_SHAPE
#undef PSC_SHAPE_Y
#undef PSC_SHAPE_PAD
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// Indices of the shapes available for (precision, N), default first; with
// `pad`, those the dealiased stepper can run its 2N passes on.
std::vector<int> shapes_for(int precision, int N, bool pad = false) {
    std::vector<int> v;
    for (int i = 0; i < kNumShapes; ++i)
        if (kShapes[i].precision == precision && kShapes[i].N == N &&
            (pad ? kShapes[i].make_pad != nullptr : kShapes[i].make != nullptr))
            v.push_back(i);
    return v;
}

void integrator_weights(const psc_config& c, std::vector<double>& w) {
    if (c.integrator == PSC_Y4) {
        if (c.yoshida_exact) {
            const double cr = std::cbrt(2.0);
            w = {1.0 / (2.0 - cr), -cr / (2.0 - cr), 1.0 / (2.0 - cr)};
        } else {
            // pspectcosmo.h:288, digits as shipped
            w = {1.351207191959657771818, -1.702414403875838200264, 1.351207191959657771818};
        }
    } else {
        w = {1.0};
    }
}

int maxnumbins(int N) { return (int)(std::sqrt(3.0) * (N / 2)) + 2; }
constexpr int kMaxLogBins = 1024;

int check_member(psc_ctx* c, int member) {
    if (!c) return PSC_ERR_ARG;
    if (member < 0 || member >= c->B) return c->fail(PSC_ERR_ARG, "member index out of range");
    return PSC_OK;
}

// Copy one member's MemberState to the host.
int read_state(psc_ctx* c, int member, MemberState* s) {
    HIPCHK(c, hipMemcpyAsync(s, c->ms + member, sizeof(MemberState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSC_OK;
}
int write_state(psc_ctx* c, int member, const MemberState* s) {
    HIPCHK(c, hipMemcpyAsync(c->ms + member, s, sizeof(MemberState), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSC_OK;
}

// FP64 half spectra (nflds fields) -> packed full spectrum (phi + i chi).
void pack_full(int N, int nflds, const double* hk, std::vector<double>& full) {
    const int Nh = N / 2 + 1;
    const size_t nk = (size_t)N * N * Nh, N3 = (size_t)N * N * N;
    full.assign(2 * N3, 0.0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                double re[2] = {0, 0}, im[2] = {0, 0};
                for (int q = 0; q < nflds; ++q) {
                    const double* h = hk + 2 * nk * q;
                    if (k <= N / 2) {
                        const size_t at = ((size_t)i * N + j) * Nh + k;
                        re[q] = h[2 * at];
                        im[q] = h[2 * at + 1];
                    } else {
                        const size_t at = ((size_t)((N - i) % N) * N + (N - j) % N) * Nh + (N - k);
                        re[q] = h[2 * at];
                        im[q] = -h[2 * at + 1];
                    }
                }
                const size_t o = ((size_t)i * N + j) * N + k;
                full[2 * o] = re[0] - im[1];
                full[2 * o + 1] = im[0] + re[1];
            }
}

// Packed full spectrum -> FP64 half spectra of nflds fields.
void unpack_half(int N, int nflds, const double* full, double* hk) {
    const int Nh = N / 2 + 1;
    const size_t nk = (size_t)N * N * Nh;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < Nh; ++k) {
                const size_t o = ((size_t)i * N + j) * N + k;
                const size_t mo = ((size_t)((N - i) % N) * N + (N - j) % N) * N + (N - k) % N;
                const double ar = full[2 * o], ai = full[2 * o + 1];
                const double br = full[2 * mo], bi = -full[2 * mo + 1];
                const size_t at = ((size_t)i * N + j) * Nh + k;
                hk[2 * at] = 0.5 * (ar + br);
                hk[2 * at + 1] = 0.5 * (ai + bi);
                if (nflds > 1) {
                    hk[2 * nk + 2 * at] = 0.5 * (ai - bi);
                    hk[2 * nk + 2 * at + 1] = -0.5 * (ar - br);
                }
            }
}

// Every copy and fill runs on the context's stream and is waited for there:
// that stream is non-blocking, so work on the null stream (hipMemcpy,
// hipMemset) would not be ordered with its kernels and could overlap them.

// Copy one unpadded N^3 complex FP64 array into slot `member` of a padded
// device array (x-planes N (N + 1) apart), converting to the field precision.
int upload(psc_ctx* c, void* dst_base, int member, const std::vector<double>& full) {
    const size_t n = c->N3, plane = (size_t)c->N * c->N;
    char* dst = static_cast<char*>(dst_base) + (size_t)member * c->MS * c->es;
    const size_t dpitch = plane_stride(c->N) * c->es, width = plane * c->es;
    if (c->cfg.precision == PSC_FP32) {
        std::vector<float> tmp(2 * n);
        for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (float)full[i];
        HIPCHK(c, hipMemcpy2DAsync(dst, dpitch, tmp.data(), width, width, c->N, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        HIPCHK(c, hipMemcpy2DAsync(dst, dpitch, full.data(), width, width, c->N, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return PSC_OK;
}

int download(psc_ctx* c, const void* src_base, int member, std::vector<double>& full) {
    const size_t n = c->N3, plane = (size_t)c->N * c->N;
    const char* src = static_cast<const char*>(src_base) + (size_t)member * c->MS * c->es;
    const size_t spitch = plane_stride(c->N) * c->es, width = plane * c->es;
    full.resize(2 * n);
    if (c->cfg.precision == PSC_FP32) {
        std::vector<float> tmp(2 * n);
        HIPCHK(c, hipMemcpy2DAsync(tmp.data(), width, src, spitch, width, c->N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < 2 * n; ++i) full[i] = tmp[i];
    } else {
        HIPCHK(c, hipMemcpy2DAsync(full.data(), width, src, spitch, width, c->N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return PSC_OK;
}

struct Diag {
    double K[2], G[2], Vt[2], mean[2], var[2], sumsq[2];
};

// Energies of one member from its device state (energy_density,
// output.cpp:57-115, with kinetic/gradient from k space and V on the lattice).
// The ghosts of the dealiased stepper as the anti-Hermitian parts they add to
// the half spectra of each field in the kz = 0, N/2 planes:
// A_phi(n) = (G(n) - conj G(-n)) / 2, A_chi(n) = (G(n) + conj G(-n)) / 2i;
// fn(field, x, y, plane, A) is called for every mode of both planes.
template <class F>
int for_each_ghost(psc_ctx* c, const void* Gsrc, F&& fn) {
    const int N = c->N;
    const size_t n = (size_t)2 * N * N;
    std::vector<double> g(2 * n);
    if (c->cfg.precision == PSC_FP32) {
        std::vector<float> tmp(2 * n);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), Gsrc, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < 2 * n; ++i) g[i] = tmp[i];
    } else {
        HIPCHK(c, hipMemcpyAsync(g.data(), Gsrc, g.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    for (int pl = 0; pl < 2; ++pl)
        for (int x = 0; x < N; ++x)
            for (int y = 0; y < N; ++y) {
                const size_t i = (size_t)pl * N * N + (size_t)x * N + y;
                const size_t m = (size_t)pl * N * N + (size_t)((N - x) % N) * N + (N - y) % N;
                const double gr = g[2 * i], gi = g[2 * i + 1], mr = g[2 * m], mi = -g[2 * m + 1];
                const double ap[2] = {0.5 * (gr - mr), 0.5 * (gi - mi)};                  // (G - conj G(-n)) / 2
                const double ac[2] = {0.5 * (gi + mi), -0.5 * (gr + mr)};                 // (G + conj G(-n)) / 2i
                fn(0, x, y, pl, ap);
                if (c->cfg.nflds > 1) fn(1, x, y, pl, ac);
            }
    return PSC_OK;
}

int diagnose(psc_ctx* c, int member, const MemberState& s, Diag* d) {
    int rc = c->eng->field_sums(c, member);
    if (rc) return rc;
    double fs[4];
    HIPCHK(c, hipMemcpyAsync(fs, c->sums, sizeof(fs), hipMemcpyDeviceToHost, c->stream));
    if ((rc = c->eng->real_space(c, member, c->Pf))) return rc;
    double rs[6];
    HIPCHK(c, hipMemcpyAsync(rs, c->sums, sizeof(rs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double N3 = (double)c->N3, a = s.a, kIR = c->cfg.kIR;
    for (int q = 0; q < 2; ++q) {
        d->K[q] = 0.5 * fs[q] / std::pow(a, 6) / (N3 * N3);
        d->G[q] = kIR * kIR * fs[2 + q] / N3 / 2.0 / (a * a) / N3;
        d->mean[q] = rs[q] / N3;
        d->sumsq[q] = rs[2 + q] / N3;
        const double v = d->sumsq[q] - d->mean[q] * d->mean[q];
        d->var[q] = v > 0 ? v : 0.0;
        d->Vt[q] = rs[4 + q] / N3;
    }
    if (c->cfg.nflds == 1) d->K[1] = d->G[1] = d->Vt[1] = d->mean[1] = d->var[1] = d->sumsq[1] = 0.0;
    if (c->padeng) {   // the half-spectrum sums include the ghosts (|H + A|^2 = |H|^2 + |A|^2 per plane)
        const int Nn = c->N;
        double gk[2] = {0, 0}, gg[2] = {0, 0};
        auto n2 = [&](int x, int y, int pl) {
            const int px = x <= Nn / 2 ? x : x - Nn, py = y <= Nn / 2 ? y : y - Nn, pz = pl ? Nn / 2 : 0;
            return (double)(px * px + py * py + pz * pz);
        };
        int rc2 = for_each_ghost(c, c->Gp, [&](int q, int, int, int, const double* A) {
            gk[q] += A[0] * A[0] + A[1] * A[1];
        });
        if (rc2) return rc2;
        rc2 = for_each_ghost(c, c->Gf, [&](int q, int x, int y, int pl, const double* A) {
            gg[q] += n2(x, y, pl) * (A[0] * A[0] + A[1] * A[1]);
        });
        if (rc2) return rc2;
        for (int q = 0; q < c->cfg.nflds; ++q) {
            d->K[q] += 0.5 * gk[q] / std::pow(a, 6) / (N3 * N3);
            d->G[q] += kIR * kIR * gg[q] / N3 / 2.0 / (a * a) / N3;
        }
    }
    return PSC_OK;
}

double kernel_a_value(const psc_ctx* c, const Diag& d, double a) {
    if (c->cfg.expansion == PSC_EXPANSION_STATIC) return 0.0;
    const int al = c->cfg.alpha;
    const double f2m = (c->cfg.fStar / c->cfg.MPl) * (c->cfg.fStar / c->cfg.MPl);
    const double K = d.K[0] + d.K[1], G = d.G[0] + d.G[1], V = d.Vt[0] + d.Vt[1];
    return 1.0 / 3.0 * f2m * std::pow(a, 1 + 2 * al) * ((al - 2) * K + al * G + (al + 1) * V);
}

// New data for the dealiased stepper has no plane ghosts: clear them before
// anything reads the state (diagnose adds them to K and G).
int reset_ghosts(psc_ctx* c) {
    if (!c->padeng) return PSC_OK;
    HIPCHK(c, hipMemsetAsync(c->Gf, 0, (size_t)2 * c->N * c->N * c->es, c->stream));
    HIPCHK(c, hipMemsetAsync(c->Gp, 0, (size_t)2 * c->N * c->N * c->es, c->stream));
    return PSC_OK;
}

// d2a/dt2 of newly loaded data (initialize.cpp:782-797): kernel_a_k_nopad, or
// kernel_a_k_pad with V averaged over the padded lattice.
int initial_add(psc_ctx* c, const Diag& d, double a, double* add) {
    Diag e = d;
    if (c->padeng) {
        double V;
        int rc = c->padeng->mean_V(c, &V);
        if (rc) return rc;
        e.Vt[0] = V;
        e.Vt[1] = 0.0;
    }
    *add = kernel_a_value(c, e, a);
    return PSC_OK;
}

// <chi^2> objective: fbar = (P(k) - conj P(-k)) / N^6 on the padded layout.
template <class G>
__global__ void k_adj_init_chi2(const cpx<typename G::T>* __restrict__ Pf, cpx<typename G::T>* __restrict__ fbar) {
    constexpr int N = G::N;
    const size_t li = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= G::N3) return;
    const int i = (int)(li / ((size_t)N * N)), j = (int)((li / N) % N), k = (int)(li % N);
    const size_t idx = (size_t)i * G::PS + (size_t)j * N + k;
    const size_t midx = (size_t)((N - i) % N) * G::PS + (size_t)((N - j) % N) * N + (N - k) % N;
    const double inv = 1.0 / ((double)G::N3 * (double)G::N3);
    const cpx<typename G::T> a = Pf[idx], b = Pf[midx];
    fbar[idx] = cpx<typename G::T>{(typename G::T)(((double)a.x - b.x) * inv), (typename G::T)(((double)a.y + b.y) * inv)};
}

template <class G, class GY>
int Engine<G, GY>::adjoint(psc_ctx* c, const psc_adjoint_spec* sp, double* out, double* dchi) {
    const int M = sp->nsteps;
    const int K = std::min(sp->checkpoint_every, M);
    const int nseg = (M + K - 1) / K;
    const int S = (int)c->w.size();
    const size_t arr = G::MS * sizeof(cpx<T>);
    const bool gw = sp->objective == PSC_OBJ_GW_FRACTION;
    AdjWork* w = nullptr;
    int rc = adj_workspace(c, K, nseg, gw, &w);
    if (rc) return rc;
    MemberState s0;
    if ((rc = read_state(c, 0, &s0))) return rc;
    Diag d0;   // K, G, V at t = 0 for the adjoint of the initial d2a/dt2
    if ((rc = diagnose(c, 0, s0, &d0))) return rc;
    hipEvent_t ev[3];
    for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
    struct Events {
        hipEvent_t* e;
        ~Events() {
            for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
        }
    } guard{ev};
    auto save = [&](int sg) -> int {
        HIPCHK(c, hipMemcpyAsync(w->ck[2 * sg], c->Pf, arr, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(w->ck[2 * sg + 1], c->Pp, arr, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(w->ck_ms + sg, c->ms, sizeof(MemberState), hipMemcpyDeviceToDevice, c->stream));
        if (gw) {
            HIPCHK(c, hipMemcpyAsync(w->ck_u[sg], c->U, 3 * arr, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(w->ck_pu[sg], c->PU, 3 * arr, hipMemcpyDeviceToDevice, c->stream));
        }
        return PSC_OK;
    };
    auto restore = [&](int sg) -> int {
        HIPCHK(c, hipMemcpyAsync(c->Pf, w->ck[2 * sg], arr, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->Pp, w->ck[2 * sg + 1], arr, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->ms, w->ck_ms + sg, sizeof(MemberState), hipMemcpyDeviceToDevice, c->stream));
        if (gw) {
            HIPCHK(c, hipMemcpyAsync(c->U, w->ck_u[sg], 3 * arr, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->PU, w->ck_pu[sg], 3 * arr, hipMemcpyDeviceToDevice, c->stream));
        }
        return PSC_OK;
    };
    // forward run with a checkpoint at every segment start; the last segment
    // is recorded on the way
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (int sg = 0; sg < nseg; ++sg) {
        if ((rc = save(sg))) return rc;
        const int steps = std::min(K, M - sg * K);
        if ((rc = sg == nseg - 1 ? run_segment(c, steps, *w, false) : step(c, steps))) return rc;
    }
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    // the objective and its adjoint at the end
    double J = 0.0;
    HIPCHK(c, hipMemsetAsync(w->pbar, 0, arr, c->stream));
    if (!gw) {
        if ((rc = real_space(c, 0, c->Pf))) return rc;
        double rs[6];
        HIPCHK(c, hipMemcpyAsync(rs, c->sums, sizeof(rs), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        J = rs[3] / (double)G::N3;
        HIPCHK(c, hipMemsetAsync(w->st, 0, sizeof(AdjState), c->stream));
        hipLaunchKernelGGL((k_adj_init_chi2<G>), dim3((unsigned)((G::N3 + 255) / 256)), dim3(256), 0, c->stream,
                           P(c->Pf), P(w->fbar));
    } else {
        // J = rho_GW / rho, rho_GW = (f*/M_P)^2 Q / (4 N^6 a^6) (output.cpp:1056-1058)
        MemberState se;
        if ((rc = read_state(c, 0, &se))) return rc;
        Diag de;
        if ((rc = diagnose(c, 0, se, &de))) return rc;
        const StepConsts sc = consts_of(c);
        const double a = se.a, N3 = (double)G::N3;
        const double Kt = de.K[0] + de.K[1], Gt = de.G[0] + de.G[1], Vt = de.Vt[0] + de.Vt[1];
        const double rho = Kt + Gt + Vt;
        const double cgw = sc.f2m / (4.0 * N3 * N3 * std::pow(a, 6));
        hipLaunchKernelGGL((k_gw_objective<G>), dim3(kObjBlocks), dim3(256), 0, c->stream, P(c->PU), P(w->pubar),
                           sp->projection, cgw / rho, w->pobj);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, c->stream, w->pobj, kObjBlocks, 1, c->sums);
        // <phi^2 chi^2 / 2> from the recorded final real-space fields (dV/dg2)
        const int nsub_last = (M - (nseg - 1) * K) * S;
        hipLaunchKernelGGL((k_vg_sums<G>), dim3(256), dim3(256), 0, c->stream, P(w->r[nsub_last]), c->part);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, c->stream, c->part, 256, 1, c->sums + 1);
        double qv[2];
        HIPCHK(c, hipMemcpyAsync(qv, c->sums, sizeof(qv), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemsetAsync(w->ubar, 0, 3 * arr, c->stream));
        HIPCHK(c, hipMemsetAsync(w->fbar, 0, arr, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        J = cgw * qv[0] / rho;
        const double vg = qv[1] / N3;
        // -J / rho times the gradient of rho = K + G + V at the final point: the
        // K and G parts and V' through one pass of the adjoint kernels
        const double cr = -J / rho;
        AdjCoef co{};
        co.ck = cr / (std::pow(a, 6) * N3 * N3);
        co.cg = cr * sc.kIR * sc.kIR / (a * a * N3 * N3);
        co.cV = cr / N3;
        co.final_pass = 1;
        HIPCHK(c, hipMemcpyAsync(w->co, &co, sizeof(co), hipMemcpyHostToDevice, c->stream));
        if ((rc = backward_passes(c, *w, nsub_last, 0, false))) return rc;
        AdjState st{};
        st.g2bar = cr * vg;
        st.abar = sc.expansion_static ? 0.0 : -6.0 * J / a + cr * (-6.0 * Kt / a - 2.0 * Gt / a);
        HIPCHK(c, hipMemcpyAsync(w->st, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    // segments from the last: restore the checkpoint, replay with recording
    // (the last segment is still recorded), sweep back
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int steps = std::min(K, M - sg * K);
        if (sg < nseg - 1) {
            if ((rc = restore(sg))) return rc;
            if ((rc = run_segment(c, steps, *w, false))) return rc;
        }
        if ((rc = run_segment(c, steps, *w, true))) return rc;
    }
    // close substep 0, then the t = 0 pass: the pending force at (f_0, a_0)
    // and the initial d2a/dt2
    AdjScalarOp op = adj_op(c);
    op.mode = 1;
    hipLaunchKernelGGL(k_adj_scalars, dim3(1), dim3(256), 0, c->stream, w->st, w->co, partials(*w), w->rms, w->rms,
                       op);
    op.mode = 2 | 4;
    op.a_e = s0.a;
    op.K_e = d0.K[0] + d0.K[1];
    op.G_e = d0.G[0] + d0.G[1];
    op.V_e = d0.Vt[0] + d0.Vt[1];
    hipLaunchKernelGGL(k_adj_scalars, dim3(1), dim3(256), 0, c->stream, w->st, w->co, partials(*w), w->rms, w->rms,
                       op);
    if ((rc = backward_passes(c, *w, 0, 0, gw))) return rc;
    op.mode = 1 | 4;
    hipLaunchKernelGGL(k_adj_scalars, dim3(1), dim3(256), 0, c->stream, w->st, w->co, partials(*w), w->rms, w->rms,
                       op);
    HIPCHK(c, hipEventRecord(ev[2], c->stream));
    HIPCHK(c, hipEventSynchronize(ev[2]));
    float fwd_ms = 0, bwd_ms = 0;
    HIPCHK(c, hipEventElapsedTime(&fwd_ms, ev[0], ev[1]));
    HIPCHK(c, hipEventElapsedTime(&bwd_ms, ev[1], ev[2]));
    AdjState hst;
    HIPCHK(c, hipMemcpyAsync(&hst, w->st, sizeof(AdjState), hipMemcpyDeviceToHost, c->stream));
    cpx<T> f00;
    HIPCHK(c, hipMemcpyAsync(&f00, w->fbar, sizeof(f00), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // phi0 enters through the DC mode phi0 N^3 and da/dt(0) = a0 H(phi0),
    // dH/dphi0 = f2m V'(phi0) / (6 H)  (initialize.cpp:649-668)
    const StepConsts sc = consts_of(c);
    const double V0 = 0.5 * s0.m2 * sp->phi0 * sp->phi0 + 0.5 * s0.g2 * sp->phi0 * sp->phi0 * sp->chi0 * sp->chi0;
    const double H = std::sqrt((sp->dphi0 * sp->dphi0 / 2.0 + V0) * sc.f2m / 3.0);
    const double vphi = s0.m2 * sp->phi0 + s0.g2 * sp->phi0 * sp->chi0 * sp->chi0;
    const double phi0bar =
        (double)G::N3 * (double)f00.x + (sc.expansion_static ? 0.0 : hst.bbar * s0.a * sc.f2m * vphi / (6.0 * H));
    // restore the loaded state
    if ((rc = restore(0))) return rc;
    if (dchi) {
        // dJ/dchi(x) = Im(F^H fbar)(x); real_space() leaves F^H src / N^3 in W
        if ((rc = real_space(c, 0, w->fbar))) return rc;
        std::vector<double> full;
        if ((rc = download(c, c->W, 0, full))) return rc;
        for (size_t i = 0; i < G::N3; ++i) dchi[i] = full[2 * i + 1] * (double)G::N3;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = J;
    out[1] = hst.g2bar;
    out[2] = phi0bar;
    out[3] = fwd_ms;
    out[4] = bwd_ms;
    out[5] = w->bytes;
    return PSC_OK;
}

}  // namespace

namespace {

std::string tune_cache_path() {
    const char* e = std::getenv("PSC_TUNE_CACHE");
    if (e && *e) return e;
    const char* h = std::getenv("HOME");
    return std::string(h ? h : ".") + "/.cache/pspectcosmo_amd/tune.txt";
}

std::string tune_key(const psc_ctx* c) {
    hipDeviceProp_t prop;
    int dev = 0;
    std::string arch = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
    for (auto& ch : arch)
        if (ch == ' ' || ch == ':') ch = '_';
    std::ostringstream k;
    k << arch << "|N" << c->N << "|" << (c->cfg.precision == PSC_FP32 ? "fp32" : "fp64") << "|B" << c->B
      << "|gw" << c->cfg.with_gw;
    if (c->cfg.pad) k << "|pad" << c->cfg.pad;
    return k.str();
}

// Last cached shape name for the key, or "".
std::string tune_lookup(const std::string& key) {
    std::ifstream in(tune_cache_path());
    std::string line, best;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string k, name;
        if (ls >> k >> name && k == key) best = name;
    }
    return best;
}

void tune_store(const std::string& key, const std::string& name, double ms) {
    const std::string path = tune_cache_path();
    for (size_t pos = path.find('/', 1); pos != std::string::npos; pos = path.find('/', pos + 1))
        mkdir(path.substr(0, pos).c_str(), 0755);
    std::ofstream out(path, std::ios::app);
    if (out) out << key << " " << name << " " << ms << "\n";
}

void drop_graphs(psc_ctx* c) {
    if (c->g_mid) (void)hipGraphExecDestroy(c->g_mid);
    if (c->g_last) (void)hipGraphExecDestroy(c->g_last);
    c->g_mid = c->g_last = nullptr;
    c->adj.reset();
}

}  // namespace

// =============================================================================
extern "C" {

int psc_abi_version(void) { return PSC_ABI_VERSION; }

const char* psc_status_string(int s) {
    switch (s) {
        case PSC_OK: return "ok";
        case PSC_ERR_ARG: return "invalid argument";
        case PSC_ERR_DEVICE: return "device error";
        case PSC_ERR_NOMEM: return "out of memory";
        case PSC_ERR_STATE: return "invalid state";
        case PSC_ERR_NONFINITE: return "field not numerical (unstable solution)";
        case PSC_ERR_NEGATIVE_A: return "scale factor is negative";
        case PSC_ERR_UNSUPPORTED: return "unsupported configuration";
        default: return "unknown status";
    }
}

int psc_num_bins(int N) { return maxnumbins(N) - 1; }

int psc_create(const psc_config* cfg, psc_ctx** out) {
    if (!cfg || !out) return PSC_ERR_ARG;
    *out = nullptr;
    const int N = cfg->N;
    if (N < 16 || N > 512 || (N & (N - 1)) || cfg->nflds < 1 || cfg->nflds > 2 || cfg->batch < 1 ||
        (cfg->integrator != PSC_VV2 && cfg->integrator != PSC_Y4) ||
        (cfg->precision != PSC_FP64 && cfg->precision != PSC_FP32) || !(cfg->kIR > 0) ||
        !(cfg->fStar > 0) || !(cfg->MPl > 0) || !std::isfinite(cfg->dt) ||
        (cfg->expansion != PSC_EXPANSION_SELF && cfg->expansion != PSC_EXPANSION_STATIC) ||
        (cfg->pad != 0 && cfg->pad != 2))
        return PSC_ERR_ARG;
    // the dealiased stepper: one member, no GW (the reference pads only the
    // scalar force), N_pad = 2N within the transform sizes
    if (cfg->pad == 2 && (cfg->batch != 1 || cfg->with_gw || N > 256)) return PSC_ERR_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PSC_ERR_DEVICE;
    std::unique_ptr<psc_ctx> c(new psc_ctx());
    c->cfg = *cfg;
    c->N = N;
    c->B = cfg->batch;
    c->N3 = (size_t)N * N * N;
    c->MS = member_stride(N);
    c->es = cfg->precision == PSC_FP32 ? 2 * sizeof(float) : 2 * sizeof(double);
    {   // ensemble groups of about 128 MiB of W each (measured on 64 x 128^3:
        // groups of 8 members FP32, 4 FP64, 8-10 % faster than one launch over
        // all members); members larger than 64 MiB run all at once
        const double member_bytes = (double)c->MS * c->es;
        const int g = (int)std::lround(134217728.0 / member_bytes);
        c->member_group = g >= 2 ? g : 0;
    }
    {
        const std::vector<int> cand = shapes_for(cfg->precision, N);
        if (cand.empty()) return PSC_ERR_UNSUPPORTED;
        c->shape = cand[0];
        if (cfg->autotune) {
            const std::string name = tune_lookup(tune_key(c.get()));
            for (int idx : cand)
                if (name == kShapes[idx].name) c->shape = idx;
        }
        c->eng = kShapes[c->shape].make();
        if (cfg->pad == 2) {
            const std::vector<int> pc = shapes_for(cfg->precision, 2 * N, true);
            if (pc.empty()) return PSC_ERR_UNSUPPORTED;
            c->pad_shape = pc[0];
            if (cfg->autotune) {
                const std::string name = tune_lookup(tune_key(c.get()));
                for (int idx : pc)
                    if (name == kShapes[idx].name) c->pad_shape = idx;
            }
            c->padeng = kShapes[c->pad_shape].make_pad();
            if (!c->padeng) return PSC_ERR_UNSUPPORTED;
        }
    }
    integrator_weights(*cfg, c->w);
    c->models.assign(c->B, psc_model{1.0, 0.0});
    c->loaded.assign(c->B, 0);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return PSC_ERR_DEVICE;
    const size_t arr = (size_t)c->B * c->MS * c->es;
    auto alloc = [&](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        c->dev_bytes += (double)bytes;
        return true;
    };
    // partial-sum sizes: the most blocks any shape of the table can launch
    // (TC >= 8, L >= 4, asserted in Shape), also after a shape switch; the
    // dealiased stepper's x pass runs on its 2N grid
    size_t blocks_col = (size_t)(N / 8) * N;
    const size_t blocks_row = (size_t)N * N / 4;
    if (c->padeng) blocks_col = std::max(blocks_col, (size_t)(2 * N / 8) * (2 * N));
    c->part_len = std::max<size_t>(std::max(blocks_col * 6, (size_t)256 * 4),
                                   (size_t)128 * (maxnumbins(N) * 4 + kMaxLogBins));
    c->sums_len = (size_t)maxnumbins(N) * 4 + kMaxLogBins + 64;
    const bool gw = cfg->with_gw != 0;
    if ((gw && (!alloc(&c->U, 3 * arr) || !alloc(&c->PU, 3 * arr))) ||
        !alloc(&c->Pf, arr) || !alloc(&c->Pp, arr) || !alloc(&c->W, gw ? 4 * arr : arr) ||
        !alloc(&c->tw, (size_t)N * c->es) || !alloc((void**)&c->ms, sizeof(MemberState) * c->B) ||
        !alloc((void**)&c->partV, sizeof(double) * blocks_col * c->B) ||
        !alloc((void**)&c->partKG, sizeof(double) * 2 * blocks_row * c->B) ||
        !alloc((void**)&c->partV2, sizeof(double) * kPreChunks * c->B) ||
        !alloc((void**)&c->partKG2, sizeof(double) * 2 * kPreChunks * c->B) ||
        !alloc((void**)&c->part, sizeof(double) * c->part_len) ||
        !alloc((void**)&c->sums, sizeof(double) * c->sums_len) ||
        (c->padeng && (!alloc(&c->Wp, c->padeng->padded_elems() * c->es) ||
                       !alloc(&c->Rp, (size_t)(N + 1) * (N + 1) * (N + 1) * c->es) ||
                       !alloc(&c->Gf, (size_t)2 * N * N * c->es) || !alloc(&c->Gp, (size_t)2 * N * N * c->es) ||
                       !alloc(&c->dcp, c->es) || !alloc(&c->twp, (size_t)2 * N * c->es)))) {
        psc_destroy(c.release());
        return PSC_ERR_NOMEM;
    }
    // Twiddles exp(-2 pi i m / n), m < n, rounded once from long double.
    auto twiddles = [&](void* dst, int n) {
        std::vector<double> twd(2 * n);
        for (int m = 0; m < n; ++m) {
            long double ang = -2.0L * 3.141592653589793238462643383279502884L * m / n;
            twd[2 * m] = (double)cosl(ang);
            twd[2 * m + 1] = (double)sinl(ang);
        }
        if (cfg->precision == PSC_FP32) {
            std::vector<float> twf(2 * n);
            for (int i = 0; i < 2 * n; ++i) twf[i] = (float)twd[i];
            return hipMemcpyAsync(dst, twf.data(), twf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream) ==
                       hipSuccess &&
                   hipStreamSynchronize(c->stream) == hipSuccess;
        }
        return hipMemcpyAsync(dst, twd.data(), twd.size() * sizeof(double), hipMemcpyHostToDevice, c->stream) ==
                   hipSuccess &&
               hipStreamSynchronize(c->stream) == hipSuccess;
    };
    if (!twiddles(c->tw, N) || (c->padeng && !twiddles(c->twp, 2 * N))) {
        psc_destroy(c.release());
        return PSC_ERR_DEVICE;
    }
    std::vector<MemberState> init(c->B);
    for (auto& s : init) {
        std::memset(&s, 0, sizeof(s));
        s.a = s.a_start = s.a_half = 1.0;
        s.m2 = 1.0;
    }
    // the fills finish before the context is handed out: they are on the
    // context's stream, ordered before anything the caller launches there
    auto fill = [&](void* p, size_t bytes) { return hipMemsetAsync(p, 0, bytes, c->stream) == hipSuccess; };
    if (hipMemcpyAsync(c->ms, init.data(), sizeof(MemberState) * c->B, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        !fill(c->Pf, arr) || !fill(c->Pp, arr) || !fill(c->W, gw ? 4 * arr : arr) ||
        (c->padeng && (!fill(c->Gf, (size_t)2 * N * N * c->es) || !fill(c->Gp, (size_t)2 * N * N * c->es))) ||
        (gw && (!fill(c->U, 3 * arr) || !fill(c->PU, 3 * arr))) || hipStreamSynchronize(c->stream) != hipSuccess) {
        psc_destroy(c.release());
        return PSC_ERR_DEVICE;
    }
    *out = c.release();
    return PSC_OK;
}

int psc_destroy(psc_ctx* c) {
    if (!c) return PSC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (auto& p : c->pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (auto e : c->ev_free) hipEventDestroy(e);
    void* ptrs[] = {c->Pf, c->Pp, c->W, c->tw, c->U, c->PU, c->ms, c->partV, c->partKG, c->partV2, c->partKG2,
                    c->part, c->sums, c->Wp, c->Rp, c->Gf, c->Gp, c->dcp, c->twp};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c->eng;
    delete c->padeng;
    delete c;
    return PSC_OK;
}

const char* psc_last_error(const psc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int psc_set_model(psc_ctx* c, int member, const psc_model* m) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!m) return c->fail(PSC_ERR_ARG, "null model");
    c->models[member] = *m;
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    s.m2 = m->m2;
    s.g2 = m->g2;
    return write_state(c, member, &s);
}

int psc_load_ic_k(psc_ctx* c, int member, const double* f_k, const double* pi_k, double a, double ad) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!f_k || !pi_k || !(a > 0)) return c->fail(PSC_ERR_ARG, "bad initial data");
    std::vector<double> full;
    pack_full(c->N, c->cfg.nflds, f_k, full);
    if ((rc = upload(c, c->Pf, member, full))) return rc;
    pack_full(c->N, c->cfg.nflds, pi_k, full);
    if ((rc = upload(c, c->Pp, member, full))) return rc;
    if ((rc = reset_ghosts(c))) return rc;
    MemberState s;
    std::memset(&s, 0, sizeof(s));
    s.t = 0.0;
    s.a = s.a_start = s.a_half = a;
    s.ad = s.b = (c->cfg.expansion == PSC_EXPANSION_STATIC) ? 0.0 : ad;
    s.m2 = c->models[member].m2;
    s.g2 = c->models[member].g2;
    if ((rc = write_state(c, member, &s))) return rc;
    Diag d;
    if ((rc = diagnose(c, member, s, &d))) return rc;
    if ((rc = initial_add(c, d, a, &s.add))) return rc;  // initialize.cpp:797
    if ((rc = write_state(c, member, &s))) return rc;
    c->loaded[member] = 1;
    return PSC_OK;
}

namespace {

// Constants of one device generation call: psc_generate_ic's background and
// vacuum normalisation (initialize.cpp:624-668, 685-700) or, for the GW
// modes, initialize_GW's (initialize.cpp:341-360).  sc = a, da/dt, printed
// d2a/dt2, H.
int ic_setup(psc_ctx* c, int member, const psc_ic_params* ic, bool gw, IcConsts* k, double sc[4]) {
#pragma clang fp contract(off)
    const psc_config& cfg = c->cfg;
    const psc_model& mdl = c->models[member];
    const psc_ic::Pot V{cfg.nflds, mdl.m2, mdl.g2};
    double a, ad, addp, H;
    psc_ic::background(cfg, V, *ic, &a, &ad, &addp, &H);
    if (!(H >= 0.0)) return c->fail(PSC_ERR_ARG, "negative H^2 from the initial values");
    const int N = c->N;
    const double L = 2.0 * 3.14159265358979323846264338327950288 / cfg.kIR, dx = L / N;
    std::memset(k, 0, sizeof(*k));
    k->N = N;
    k->nflds = cfg.nflds;
    k->seed = ic->seed;
    k->a = a;
    k->ap = std::pow(a, cfg.alpha - 3);
    k->H = H;
    k->N3 = (double)N * N * N;
    k->kIR = cfg.kIR;
    k->dp2 = cfg.kIR * cfg.kIR;
    const unsigned long long M = psc_ic::mode_count(N, true);
    if (!gw) {
        k->norm = cfg.omegaStar / cfg.fStar * std::pow(N / dx, 1.5) / std::sqrt(2.0);
        const double f0[2] = {ic->f0[0], cfg.nflds > 1 ? ic->f0[1] : 0.0};
        V.meff(k->meff, f0);
        for (int q = 0; q < cfg.nflds; ++q) {
            k->first[q] = 4 * (unsigned long long)q * M;
            k->f0[q] = ic->f0[q];
            k->fd0[q] = ic->fd0[q];
        }
    } else {
        k->norm = cfg.MPl * cfg.omegaStar / (cfg.fStar * cfg.fStar) * std::pow(N / dx, 1.5);
        k->first[0] = 4 * M * (unsigned long long)cfg.nflds;   // after every scalar mode's draws
    }
    sc[0] = a;
    sc[1] = ad;
    sc[2] = addp;
    sc[3] = H;
    return PSC_OK;
}

}  // namespace

int psc_generate_ic_device(psc_ctx* c, int member, const psc_ic_params* ic, double* scalars_out) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!ic || !(ic->a0 > 0)) return c->fail(PSC_ERR_ARG, "bad initial-condition parameters");
    IcConsts k;
    double sc[4];
    if ((rc = ic_setup(c, member, ic, false, &k, sc))) return rc;
    const unsigned long long count = psc_ic::mode_count(c->N, true);
    const dim3 grid((unsigned)((count + 255) / 256)), block(256);
    const size_t off = (size_t)member * c->MS, PS = plane_stride(c->N);
    if (c->cfg.precision == PSC_FP32)
        hipLaunchKernelGGL(k_ic_fields<float>, grid, block, 0, c->stream, static_cast<cpx<float>*>(c->Pf) + off,
                           static_cast<cpx<float>*>(c->Pp) + off, k, PS, count);
    else
        hipLaunchKernelGGL(k_ic_fields<double>, grid, block, 0, c->stream, static_cast<cpx<double>*>(c->Pf) + off,
                           static_cast<cpx<double>*>(c->Pp) + off, k, PS, count);
    HIPCHK(c, hipGetLastError());
    if ((rc = reset_ghosts(c))) return rc;
    // the member's state as psc_load_ic_k sets it
    MemberState s;
    std::memset(&s, 0, sizeof(s));
    s.a = s.a_start = s.a_half = sc[0];
    s.ad = s.b = (c->cfg.expansion == PSC_EXPANSION_STATIC) ? 0.0 : sc[1];
    s.m2 = c->models[member].m2;
    s.g2 = c->models[member].g2;
    if ((rc = write_state(c, member, &s))) return rc;
    Diag d;
    if ((rc = diagnose(c, member, s, &d))) return rc;
    if ((rc = initial_add(c, d, sc[0], &s.add))) return rc;  // initialize.cpp:797
    if ((rc = write_state(c, member, &s))) return rc;
    c->loaded[member] = 1;
    if (scalars_out)
        for (int i = 0; i < 4; ++i) scalars_out[i] = sc[i];
    return PSC_OK;
}

int psc_generate_gw_ic_device(psc_ctx* c, int member, const psc_ic_params* ic) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!c->cfg.with_gw) return c->fail(PSC_ERR_STATE, "context created without GW evolution");
    if (!ic || !(ic->a0 > 0)) return c->fail(PSC_ERR_ARG, "bad initial-condition parameters");
    IcConsts k;
    double sc[4];
    if ((rc = ic_setup(c, member, ic, true, &k, sc))) return rc;
    const size_t arr = c->MS * c->es, astride = (size_t)c->B * c->MS;
    for (int a = 0; a < 3; ++a) {   // the DC mode is not generated and stays zero
        const size_t o = ((size_t)a * c->B + member) * arr;
        HIPCHK(c, hipMemsetAsync(static_cast<char*>(c->U) + o, 0, arr, c->stream));
        HIPCHK(c, hipMemsetAsync(static_cast<char*>(c->PU) + o, 0, arr, c->stream));
    }
    const unsigned long long count = psc_ic::mode_count(c->N, false);
    const dim3 grid((unsigned)((count + 255) / 256)), block(256);
    const size_t off = (size_t)member * c->MS, PS = plane_stride(c->N);
    if (c->cfg.precision == PSC_FP32)
        hipLaunchKernelGGL(k_ic_gw<float>, grid, block, 0, c->stream, static_cast<cpx<float>*>(c->U) + off,
                           static_cast<cpx<float>*>(c->PU) + off, astride, k, PS, count);
    else
        hipLaunchKernelGGL(k_ic_gw<double>, grid, block, 0, c->stream, static_cast<cpx<double>*>(c->U) + off,
                           static_cast<cpx<double>*>(c->PU) + off, astride, k, PS, count);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->loaded_gw = 1;
    return PSC_OK;
}

int psc_load_fields(psc_ctx* c, int member, const double* f, const double* pi, double t, double a, double ad) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!f || !pi || !(a > 0)) return c->fail(PSC_ERR_ARG, "bad lattice image");
    const size_t n3 = c->N3;
    std::vector<double> full(2 * n3);
    for (int which = 0; which < 2; ++which) {
        const double* src = which == 0 ? f : pi;
        for (size_t i = 0; i < n3; ++i) {
            full[2 * i] = src[i];
            full[2 * i + 1] = c->cfg.nflds > 1 ? src[n3 + i] : 0.0;
        }
        if ((rc = upload(c, c->W, member, full))) return rc;
        if ((rc = c->eng->to_spectrum(c, member, which == 0 ? c->Pf : c->Pp))) return rc;
    }
    if ((rc = reset_ghosts(c))) return rc;
    MemberState s;
    std::memset(&s, 0, sizeof(s));
    s.t = t;
    s.a = s.a_start = s.a_half = a;
    s.ad = s.b = (c->cfg.expansion == PSC_EXPANSION_STATIC) ? 0.0 : ad;
    s.m2 = c->models[member].m2;
    s.g2 = c->models[member].g2;
    if ((rc = write_state(c, member, &s))) return rc;
    Diag d;
    if ((rc = diagnose(c, member, s, &d))) return rc;
    if ((rc = initial_add(c, d, a, &s.add))) return rc;  // initialize.cpp:797, as for new initial data
    if ((rc = write_state(c, member, &s))) return rc;
    c->loaded[member] = 1;
    return PSC_OK;
}

int psc_set_scalars(psc_ctx* c, int member, double t, double a, double ad, double add) {
    int rc = check_member(c, member);
    if (rc) return rc;
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    s.t = t;
    s.a = s.a_start = s.a_half = a;
    s.ad = s.b = ad;
    s.add = add;
    return write_state(c, member, &s);
}

int psc_get_scalars(psc_ctx* c, int member, double* out) {
    int rc = check_member(c, member);
    if (rc) return rc;
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    out[0] = s.t;
    out[1] = s.a;
    out[2] = s.ad;
    out[3] = s.add;
    if (s.a < 0) return c->fail(PSC_ERR_NEGATIVE_A, "scale factor is negative");
    return PSC_OK;
}

int psc_step(psc_ctx* c, int nsteps) {
    if (!c) return PSC_ERR_ARG;
    if (nsteps < 0) return c->fail(PSC_ERR_ARG, "negative step count");
    for (int b = 0; b < c->B; ++b)
        if (!c->loaded[b]) return c->fail(PSC_ERR_STATE, "member " + std::to_string(b) + " has no initial data");
    int rc = c->eng->step(c, nsteps);
    if (rc) return rc;
    HIPCHK(c, hipGetLastError());
    return PSC_OK;
}

int psc_sync(psc_ctx* c) {
    if (!c) return PSC_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSC_OK;
}

int psc_get_k(psc_ctx* c, int member, double* f_k, double* pi_k) {
    int rc = check_member(c, member);
    if (rc) return rc;
    std::vector<double> full;
    const size_t nk = (size_t)c->N * c->N * (c->N / 2 + 1);
    auto add_ghost = [&](const void* G, double* out) {   // the dealiased stepper's plane ghosts
        return for_each_ghost(c, G, [&](int q, int x, int y, int pl, const double* A) {
            const size_t i = 2 * (q * nk + ((size_t)x * c->N + y) * (c->N / 2 + 1) + (pl ? c->N / 2 : 0));
            out[i] += A[0];
            out[i + 1] += A[1];
        });
    };
    if (f_k) {
        if ((rc = download(c, c->Pf, member, full))) return rc;
        unpack_half(c->N, c->cfg.nflds, full.data(), f_k);
        if (c->padeng && (rc = add_ghost(c->Gf, f_k))) return rc;
    }
    if (pi_k) {
        if ((rc = download(c, c->Pp, member, full))) return rc;
        unpack_half(c->N, c->cfg.nflds, full.data(), pi_k);
        if (c->padeng && (rc = add_ghost(c->Gp, pi_k))) return rc;
    }
    return PSC_OK;
}

int psc_get_fields(psc_ctx* c, int member, double* f, double* pi) {
    int rc = check_member(c, member);
    if (rc) return rc;
    std::vector<double> full;
    for (int which = 0; which < 2; ++which) {
        double* out = which == 0 ? f : pi;
        if (!out) continue;
        if ((rc = c->eng->real_space(c, member, which == 0 ? c->Pf : c->Pp))) return rc;
        if ((rc = download(c, c->W, member, full))) return rc;
        for (size_t i = 0; i < c->N3; ++i) {
            out[i] = full[2 * i];
            if (c->cfg.nflds > 1) out[c->N3 + i] = full[2 * i + 1];
        }
    }
    return PSC_OK;
}

int psc_energy_densities(psc_ctx* c, int member, double* pot, double* kin, double* grad, double* tot) {
    int rc = check_member(c, member);
    if (rc) return rc;
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    const size_t n3 = c->N3;
    double* D = nullptr;
    HIPCHK(c, hipMalloc(&D, n3 * sizeof(double)));
    std::unique_ptr<double, void (*)(double*)> hold(D, [](double* p) { (void)hipFree(p); });
    std::vector<double> tmp;
    double* outs[3] = {pot, kin, grad};
    // potential (coupling term only with two fields), a^6/2 sum pi^2,
    // a^2/2 sum |grad f|^2 (evolution.cpp:7-19, 135-137)
    const double scales[3] = {c->cfg.nflds == 2 ? 1.0 : 0.0, std::pow(s.a, 6) / 2.0, s.a * s.a / 2.0};
    for (int kind = 0; kind < 3; ++kind) {
        double* dst = outs[kind];
        if (!dst && !tot) continue;
        if (!dst) {
            tmp.resize(n3);
            dst = tmp.data();
        }
        if ((rc = c->eng->density(c, member, kind, D, scales[kind]))) return rc;
        HIPCHK(c, hipMemcpyAsync(dst, D, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (tot) {  // tot = pot + kin + grad in that order (output.cpp:1623)
            if (kind == 0)
                std::memcpy(tot, dst, n3 * sizeof(double));
            else
                for (size_t i = 0; i < n3; ++i) tot[i] += dst[i];
        }
    }
    return PSC_OK;
}

int psc_set_dt(psc_ctx* c, double dt) {
    if (!c) return PSC_ERR_ARG;
    if (!(dt > 0.0) || !std::isfinite(dt)) return c->fail(PSC_ERR_ARG, "dt must be positive");
    if (dt != c->cfg.dt) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->cfg.dt = dt;
        drop_graphs(c);
    }
    return PSC_OK;
}

int psc_means(psc_ctx* c, int member, double* out) {
    int rc = check_member(c, member);
    if (rc) return rc;
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    Diag d;
    if ((rc = diagnose(c, member, s, &d))) return rc;
    // Sum |pi_k|^2 per field (the kinetic sums) and the k = 0 mode of P_pi
    if ((rc = c->eng->field_sums(c, member))) return rc;
    double fs[4];
    HIPCHK(c, hipMemcpyAsync(fs, c->sums, sizeof(fs), hipMemcpyDeviceToHost, c->stream));
    double p0[2] = {0.0, 0.0};
    const char* src = static_cast<const char*>(c->Pp) + (size_t)member * c->MS * c->es;
    if (c->cfg.precision == PSC_FP32) {
        float v[2];
        HIPCHK(c, hipMemcpyAsync(v, src, sizeof(v), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        p0[0] = v[0];
        p0[1] = v[1];
    } else {
        HIPCHK(c, hipMemcpyAsync(p0, src, sizeof(p0), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    const double N3 = (double)c->N3, aop = std::pow(s.a, c->cfg.alpha - 3);
    for (int q = 0; q < c->cfg.nflds; ++q) {
        const double pav = p0[q] / N3, psq = fs[q] / (N3 * N3);
        const double vf = d.sumsq[q] - d.mean[q] * d.mean[q], vp = psq - pav * pav;
        double* o = out + 6 * q;
        o[0] = d.mean[q];
        o[1] = aop * pav;
        o[2] = d.sumsq[q];
        o[3] = aop * aop * psq;
        o[4] = vf > 0 ? std::sqrt(vf) : 0.0;
        o[5] = vp > 0 ? aop * std::sqrt(vp) : 0.0;
        if (!std::isfinite(d.mean[q]))  // output.cpp:450-454
            return c->fail(PSC_ERR_NONFINITE, "field " + std::to_string(q) + " not numerical at t = " + std::to_string(s.t));
    }
    return PSC_OK;
}

int psc_energies(psc_ctx* c, int member, double* out) {
    int rc = check_member(c, member);
    if (rc) return rc;
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    Diag d;
    if ((rc = diagnose(c, member, s, &d))) return rc;
    for (int i = 0; i < PSC_ENERGY_LEN; ++i) out[i] = 0.0;
    double total = 0.0;
    for (int q = 0; q < c->cfg.nflds; ++q) {
        out[2 * q] = d.K[q];
        out[2 * q + 1] = d.G[q];
        total += d.K[q] + d.G[q];
    }
    const int nterms = c->cfg.nflds == 2 ? 2 : 1;
    for (int t = 0; t < nterms; ++t) {
        out[4 + t] = d.Vt[t];
        total += d.Vt[t];
    }
    out[6] = total;
    const int al = c->cfg.alpha;
    const double f2m = (c->cfg.fStar / c->cfg.MPl) * (c->cfg.fStar / c->cfg.MPl);
    const double LHS = s.ad * s.ad, RHS = f2m * std::pow(s.a, 2 * (al + 1)) * total / 3.0;
    out[7] = std::fabs(LHS - RHS) / std::fabs(LHS + RHS);  // output.cpp:1180-1184
    out[8] = s.a;
    out[9] = s.ad;
    out[10] = s.add;
    out[11] = s.t;
    out[12] = d.mean[0];
    out[13] = d.mean[1];
    out[14] = d.var[0];
    out[15] = d.var[1];
    for (int q = 0; q < c->cfg.nflds; ++q)  // output.cpp:450-454
        if (!std::isfinite(d.mean[q]))
            return c->fail(PSC_ERR_NONFINITE, "field " + std::to_string(q) + " not numerical at t = " + std::to_string(s.t));
    if (s.a < 0) return c->fail(PSC_ERR_NEGATIVE_A, "a is negative at t = " + std::to_string(s.t));  // output.cpp:480-483
    return PSC_OK;
}

int psc_spectra(psc_ctx* c, int member, int fld, double* out, int nbins) {
    int rc = check_member(c, member);
    if (rc) return rc;
    const int mb = maxnumbins(c->N);
    if (fld < 0 || fld >= c->cfg.nflds || nbins != mb - 1 || !out) return c->fail(PSC_ERR_ARG, "bad spectra arguments");
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    // <m_eff^2> of the field (effective_mass, model.h:96-98, averaged as in
    // output.cpp:597-608): m2 + g2 <chi^2> for phi, g2 <phi^2> for chi.
    if ((rc = c->eng->real_space(c, member, c->Pf))) return rc;
    double rs[6];
    HIPCHK(c, hipMemcpyAsync(rs, c->sums, sizeof(rs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double N3 = (double)c->N3;
    const double msq = c->cfg.nflds == 1 ? s.m2
                                         : (fld == 0 ? s.m2 + s.g2 * rs[3] / N3 : s.g2 * rs[2] / N3);
    if ((rc = c->eng->spectra(c, member, fld))) return rc;
    std::vector<double> h((size_t)mb * 4);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double a = s.a, kIR = c->cfg.kIR, L = 2.0 * 3.14159265358979323846 / kIR;
    const double aop = std::pow(a, c->cfg.alpha - 3);
    if (c->padeng) {
        // the plane ghosts of the dealiased stepper add |A|^2 to their bins
        // (weight 1; the cross terms with the Hermitian parts cancel between
        // n and -n, which share a bin)
        const int N = c->N;
        std::vector<double> af((size_t)2 * N * N * 2);
        if ((rc = for_each_ghost(c, c->Gf, [&](int q, int x, int y, int pl, const double* A) {
                 if (q != fld) return;
                 const size_t i = ((size_t)pl * N * N + (size_t)x * N + y) * 2;
                 af[i] = A[0];
                 af[i + 1] = A[1];
             })))
            return rc;
        const double hub = s.ad / a;
        if ((rc = for_each_ghost(c, c->Gp, [&](int q, int x, int y, int pl, const double* A) {
                 if (q != fld) return;
                 const size_t i = ((size_t)pl * N * N + (size_t)x * N + y) * 2;
                 const int px = x <= N / 2 ? x : x - N, py = y <= N / 2 ? y : y - N, pz = pl ? N / 2 : 0;
                 const int l = (int)std::floor(std::sqrt((double)(px * px + py * py + pz * pz)) + 0.5);
                 const double nr = aop * A[0] + hub * af[i], ni = aop * A[1] + hub * af[i + 1];
                 h[(size_t)l * 4 + 1] += af[i] * af[i] + af[i + 1] * af[i + 1];
                 h[(size_t)l * 4 + 2] += A[0] * A[0] + A[1] * A[1];
                 h[(size_t)l * 4 + 3] += nr * nr + ni * ni;
             })))
            return rc;
    }
    for (int i = 1; i < mb; ++i) {
        const double np = h[(size_t)i * 4];
        double f2 = h[(size_t)i * 4 + 1], p2 = h[(size_t)i * 4 + 2], n2 = h[(size_t)i * 4 + 3];
        if (np > 0) {
            f2 /= np;
            p2 /= np;
            n2 /= np;
        }
        const double p = kIR * i;
        const double om2 = std::fabs(p * p + a * a * msq);
        double nd = 0.0;
        if (om2 > 0) {
            const double om = std::sqrt(om2);
            nd = a * a * std::pow(L, 3) / 2.0 / (N3 * N3) *
                 (c->cfg.fStar / c->cfg.omegaStar) * (c->cfg.fStar / c->cfg.omegaStar) *
                 (om * f2 + std::pow(a, 2 * (1 - c->cfg.alpha)) / om * n2);
        }
        double* row = out + (size_t)(i - 1) * PSC_SPECTRA_COLS;
        row[0] = p;
        row[1] = i / (N3 * N3) * np * f2;
        row[2] = i / (N3 * N3) * np * aop * aop * p2;
        row[3] = nd;
        row[4] = np;
    }
    return PSC_OK;
}

int psc_load_gw_k(psc_ctx* c, int member, const double* u_k, const double* pi_u_k) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!c->cfg.with_gw) return c->fail(PSC_ERR_STATE, "context created without GW evolution");
    if (!u_k || !pi_u_k) return c->fail(PSC_ERR_ARG, "null GW data");
    const size_t nk = (size_t)c->N * c->N * (c->N / 2 + 1);
    std::vector<double> full;
    for (int a = 0; a < 3; ++a) {
        const size_t arr_off = (size_t)a * c->B;  // pair a of member m sits at [a][m]
        pack_full(c->N, 2, u_k + 2 * nk * (2 * a), full);
        if ((rc = upload(c, c->U, (int)(arr_off + member), full))) return rc;
        pack_full(c->N, 2, pi_u_k + 2 * nk * (2 * a), full);
        if ((rc = upload(c, c->PU, (int)(arr_off + member), full))) return rc;
    }
    c->loaded_gw = 1;
    return PSC_OK;
}

int psc_get_gw_k(psc_ctx* c, int member, double* u_k, double* pi_u_k) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!c->cfg.with_gw) return c->fail(PSC_ERR_STATE, "context created without GW evolution");
    const size_t nk = (size_t)c->N * c->N * (c->N / 2 + 1);
    std::vector<double> full;
    for (int a = 0; a < 3; ++a) {
        const int slot = a * c->B + member;
        if (u_k) {
            if ((rc = download(c, c->U, slot, full))) return rc;
            unpack_half(c->N, 2, full.data(), u_k + 2 * nk * (2 * a));
        }
        if (pi_u_k) {
            if ((rc = download(c, c->PU, slot, full))) return rc;
            unpack_half(c->N, 2, full.data(), pi_u_k + 2 * nk * (2 * a));
        }
    }
    return PSC_OK;
}

namespace {
// Shared part of psc_gw_spectrum / psc_gw_omega_log: histograms, total energy
// and the DC kinetic term sum_ij |pi_u_ij(0)|^2 (output.cpp:1038-1044).
int gw_histograms(psc_ctx* c, int member, int projection, int nlog, std::vector<double>& h, double* total,
                  double* hd0, MemberState* s) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!c->cfg.with_gw) return c->fail(PSC_ERR_STATE, "context created without GW evolution");
    if (projection != PSC_GW_PROJ_REFERENCE && projection != PSC_GW_PROJ_TT)
        return c->fail(PSC_ERR_ARG, "unknown projection");
    if (nlog < 1 || nlog > kMaxLogBins) return c->fail(PSC_ERR_ARG, "log bin count out of range");
    if ((rc = read_state(c, member, s))) return rc;
    double e[PSC_ENERGY_LEN];
    rc = psc_energies(c, member, e);
    if (rc && rc != PSC_ERR_NONFINITE) return rc;
    *total = e[6];
    if ((rc = c->eng->gw_spectra(c, member, projection, nlog))) return rc;
    const int mb = maxnumbins(c->N);
    h.resize((size_t)3 * mb + nlog);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    *hd0 = 0.0;
    for (int a = 0; a < 3; ++a) {
        double dc[2];
        const char* src = static_cast<const char*>(c->PU) + ((size_t)a * c->B + member) * c->MS * c->es;
        if (c->cfg.precision == PSC_FP32) {
            float f2[2];
            HIPCHK(c, hipMemcpyAsync(f2, src, sizeof(f2), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            dc[0] = f2[0];
            dc[1] = f2[1];
        } else {
            HIPCHK(c, hipMemcpyAsync(dc, src, sizeof(dc), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        // pair a = (component 2a, component 2a+1); at k = 0 both are real:
        // u_{2a}(0) = Re P(0), u_{2a+1}(0) = Im P(0).  Off-diagonal components
        // (index >= 3) count twice in the 9-component sum.
        const double w0 = (2 * a) >= 3 ? 2.0 : 1.0, w1 = (2 * a + 1) >= 3 ? 2.0 : 1.0;
        *hd0 += w0 * dc[0] * dc[0] + w1 * dc[1] * dc[1];
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSC_OK;
}
}  // namespace

int psc_gw_spectrum(psc_ctx* c, int member, int projection, double* out, int nbins, double* totals) {
    if (!c) return PSC_ERR_ARG;
    const int mb = maxnumbins(c->N);
    if (!out || nbins != mb - 1) return c->fail(PSC_ERR_ARG, "bad GW spectrum arguments");
    std::vector<double> h;
    double total, hd0;
    MemberState s;
    int rc = gw_histograms(c, member, projection, 1, h, &total, &hd0, &s);
    if (rc) return rc;
    const double N3 = (double)c->N3, a = s.a, ad = s.ad;
    const double fS = c->cfg.fStar, wS = c->cfg.omegaStar, MP = c->cfg.MPl, kIR = c->cfg.kIR;
    const double pi = 3.14159265358979323846;
    double rho = hd0;
    for (int i = 1; i < mb; ++i) {
        const double np = h[(size_t)i * 3];
        double h2 = h[(size_t)i * 3 + 1], hd2 = h[(size_t)i * 3 + 2];
        rho += hd2;
        if (np > 0) {
            h2 /= np;
            hd2 /= np;
        }
        double* row = out + (size_t)(i - 1) * PSC_GW_SPECTRA_COLS;
        row[0] = kIR * i;
        row[1] = i / (N3 * N3) * np * h2;
        row[2] = i / (N3 * N3) / 4.0 / std::pow(a, 6) * (fS / MP) * (fS / MP) / total * np * hd2;
        row[3] = (MP * wS) * (MP * wS) / std::pow(fS, 4) * 2 / (pi * pi) * ad * ad / std::pow(a, 2 * c->cfg.alpha + 2);
        row[4] = np;
    }
    rho *= wS * wS * std::pow(fS, 4) / (MP * MP) / 4.0 / (N3 * N3) / std::pow(a, 6);
    rho /= (fS * wS) * (fS * wS);
    if (totals) {
        totals[0] = rho / total;
        totals[1] = rho;
    }
    return PSC_OK;
}

int psc_gw_omega_log(psc_ctx* c, int member, int projection, int nlog, double* k_center, double* omega) {
    if (!c) return PSC_ERR_ARG;
    if (!k_center || !omega) return c->fail(PSC_ERR_ARG, "null output");
    std::vector<double> h;
    double total, hd0;
    MemberState s;
    int rc = gw_histograms(c, member, projection, nlog, h, &total, &hd0, &s);
    if (rc) return rc;
    const int mb = maxnumbins(c->N);
    const double N3 = (double)c->N3, a = s.a;
    const double f2m = (c->cfg.fStar / c->cfg.MPl) * (c->cfg.fStar / c->cfg.MPl);
    const double lmax = std::log(std::sqrt(3.0) * (c->N / 2)), dl = lmax / nlog;
    for (int b = 0; b < nlog; ++b) {
        k_center[b] = c->cfg.kIR * std::exp((b + 0.5) * dl);
        const double rho_b = f2m / 4.0 / (N3 * N3) / std::pow(a, 6) * h[(size_t)3 * mb + b];
        omega[b] = rho_b / total / dl;
    }
    return PSC_OK;
}


int psc_autotune(psc_ctx* c, int nsteps, char* report, int report_len) {
    if (!c) return PSC_ERR_ARG;
    if (nsteps < 1) return c->fail(PSC_ERR_ARG, "nsteps must be >= 1");
    for (int b = 0; b < c->B; ++b)
        if (!c->loaded[b]) return c->fail(PSC_ERR_STATE, "autotune needs loaded initial data");
    const size_t arr = (size_t)c->B * c->MS * c->es;
    const bool gw = c->cfg.with_gw != 0;
    const bool pad = c->padeng != nullptr;
    const size_t garr = pad ? (size_t)2 * c->N * c->N * c->es : 0;
    void *sf = nullptr, *sp = nullptr, *su = nullptr, *spu = nullptr, *sgf = nullptr, *sgp = nullptr;
    std::vector<MemberState> ms(c->B);
    auto cleanup = [&]() {
        for (void* q : {sf, sp, su, spu, sgf, sgp})
            if (q) hipFree(q);
    };
    if (hipMalloc(&sf, arr) != hipSuccess || hipMalloc(&sp, arr) != hipSuccess ||
        (gw && (hipMalloc(&su, 3 * arr) != hipSuccess || hipMalloc(&spu, 3 * arr) != hipSuccess)) ||
        (pad && (hipMalloc(&sgf, garr) != hipSuccess || hipMalloc(&sgp, garr) != hipSuccess))) {
        cleanup();
        return c->fail(PSC_ERR_NOMEM, "autotune snapshot allocation failed");
    }
    auto copy = [&](bool save) -> int {
        auto d2d = [&](void* dst, const void* src, size_t bytes) {
            return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream);
        };
        HIPCHK(c, d2d(save ? sf : c->Pf, save ? c->Pf : sf, arr));
        HIPCHK(c, d2d(save ? sp : c->Pp, save ? c->Pp : sp, arr));
        if (gw) {
            HIPCHK(c, d2d(save ? su : c->U, save ? c->U : su, 3 * arr));
            HIPCHK(c, d2d(save ? spu : c->PU, save ? c->PU : spu, 3 * arr));
        }
        if (pad) {
            HIPCHK(c, d2d(save ? sgf : c->Gf, save ? c->Gf : sgf, garr));
            HIPCHK(c, d2d(save ? sgp : c->Gp, save ? c->Gp : sgp, garr));
        }
        if (save)
            HIPCHK(c, hipMemcpyAsync(ms.data(), c->ms, sizeof(MemberState) * c->B, hipMemcpyDeviceToHost, c->stream));
        else
            HIPCHK(c, hipMemcpyAsync(c->ms, ms.data(), sizeof(MemberState) * c->B, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return PSC_OK;
    };
    int rc = copy(true);
    if (rc) { cleanup(); return rc; }
    // the dealiased stepper's time is in its 2N passes: tune those
    const std::vector<int> cand = shapes_for(c->cfg.precision, pad ? 2 * c->N : c->N, pad);
    const int saved_timing = c->timing;
    c->timing = 0;
    double best_ms = 1e300;
    int best = pad ? c->pad_shape : c->shape;
    auto adopt = [&](int idx) {
        drop_graphs(c);
        if (pad) {
            delete c->padeng;
            c->padeng = kShapes[idx].make_pad();
            c->pad_shape = idx;
        } else {
            delete c->eng;
            c->eng = kShapes[idx].make();
            c->shape = idx;
        }
    };
    std::ostringstream rep;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (int idx : cand) {
        adopt(idx);
        if ((rc = copy(false))) break;
        if ((rc = c->eng->step(c, 1))) break;   // warm-up (and graph capture)
        if ((rc = copy(false))) break;
        hipEventRecord(e0, c->stream);
        if ((rc = c->eng->step(c, nsteps))) break;
        hipEventRecord(e1, c->stream);
        hipEventSynchronize(e1);
        float ms_ = 0.0f;
        hipEventElapsedTime(&ms_, e0, e1);
        rep << kShapes[idx].name << "=" << ms_ << ";";
        if (ms_ < best_ms) {
            best_ms = ms_;
            best = idx;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c->timing = saved_timing;
    adopt(best);
    int rc2 = copy(false);
    cleanup();
    if (rc) return rc;
    if (rc2) return rc2;
    tune_store(tune_key(c), kShapes[best].name, best_ms);
    c->report = rep.str();
    if (report && report_len > 0) {
        std::strncpy(report, c->report.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return PSC_OK;
}

const char* psc_shape_name(const psc_ctx* c) {
    if (!c) return "null context";
    return kShapes[c->padeng ? c->pad_shape : c->shape].name;
}

int psc_list_shapes(const psc_ctx* c, char* names, int names_len) {
    if (!c || !names || names_len < 1) return PSC_ERR_ARG;
    std::string s;
    for (int idx : shapes_for(c->cfg.precision, c->padeng ? 2 * c->N : c->N, c->padeng != nullptr)) {
        if (!s.empty()) s += ",";
        s += kShapes[idx].name;
    }
    std::strncpy(names, s.c_str(), names_len - 1);
    names[names_len - 1] = 0;
    return PSC_OK;
}

int psc_set_shape(psc_ctx* c, const char* name) {
    if (!c || !name) return PSC_ERR_ARG;
    if (c->padeng) {   // the dealiased stepper's 2N shape
        for (int idx : shapes_for(c->cfg.precision, 2 * c->N, true))
            if (std::strcmp(name, kShapes[idx].name) == 0) {
                HIPCHK(c, hipStreamSynchronize(c->stream));
                drop_graphs(c);
                delete c->padeng;
                c->pad_shape = idx;
                c->padeng = kShapes[idx].make_pad();
                return PSC_OK;
            }
        return c->fail(PSC_ERR_ARG, std::string("no launch shape named ") + name);
    }
    for (int idx : shapes_for(c->cfg.precision, c->N))
        if (std::strcmp(name, kShapes[idx].name) == 0) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            drop_graphs(c);
            delete c->eng;
            c->shape = idx;
            c->eng = kShapes[idx].make();
            return PSC_OK;
        }
    return c->fail(PSC_ERR_ARG, std::string("no launch shape named ") + name);
}

int psc_adjoint(psc_ctx* c, const psc_adjoint_spec* sp, double* out, double* dchi) {
    if (!c) return PSC_ERR_ARG;
    if (!sp || !out) return c->fail(PSC_ERR_ARG, "null adjoint arguments");
    if (sp->objective != PSC_OBJ_CHI2 && sp->objective != PSC_OBJ_GW_FRACTION)
        return c->fail(PSC_ERR_UNSUPPORTED, "unknown objective");
    if (c->cfg.nflds != 2 || c->B != 1 || sp->member != 0)
        return c->fail(PSC_ERR_UNSUPPORTED, "the adjoint needs two fields and one member");
    if (c->padeng) return c->fail(PSC_ERR_UNSUPPORTED, "the adjoint is built for the unpadded stepper");
    if (sp->objective == PSC_OBJ_GW_FRACTION && !c->cfg.with_gw)
        return c->fail(PSC_ERR_UNSUPPORTED, "the GW objective needs a context with GW evolution");
    if (sp->objective == PSC_OBJ_GW_FRACTION &&
        (sp->projection != PSC_GW_PROJ_REFERENCE && sp->projection != PSC_GW_PROJ_TT))
        return c->fail(PSC_ERR_ARG, "unknown projection");
    if (sp->objective == PSC_OBJ_GW_FRACTION && !c->loaded_gw)
        return c->fail(PSC_ERR_STATE, "no GW initial data");
    if (sp->nsteps < 1 || sp->checkpoint_every < 1) return c->fail(PSC_ERR_ARG, "bad step counts");
    if (!c->loaded[0]) return c->fail(PSC_ERR_STATE, "no initial data");
    const int saved = c->timing;
    c->timing = 0;
    int rc = c->eng->adjoint(c, sp, out, dchi);
    c->timing = saved;
    return rc;
}

int psc_set_kernel_timing(psc_ctx* c, int on) {
    if (!c) return PSC_ERR_ARG;
    c->timing = on ? 1 : 0;
    return PSC_OK;
}

int psc_reset_kernel_times(psc_ctx* c) {
    if (!c) return PSC_ERR_ARG;
    int rc = drain_timing(c);
    for (auto& k : c->kt) k = Timing();
    return rc;
}

int psc_get_kernel_times(psc_ctx* c, char* names, int names_len, double* total_ms, long long* counts,
                         double* bytes, int cap, int* nkinds) {
    if (!c) return PSC_ERR_ARG;
    int rc = drain_timing(c);
    if (rc) return rc;
    std::string s;
    int n = 0;
    for (int k = 0; k < K_NKINDS && n < cap; ++k, ++n) {
        if (k) s += ",";
        s += kKindNames[k];
        if (total_ms) total_ms[n] = c->kt[k].ms;
        if (counts) counts[n] = c->kt[k].count;
        if (bytes) bytes[n] = c->kt[k].bytes;
    }
    if (names && names_len > 0) {
        std::strncpy(names, s.c_str(), names_len - 1);
        names[names_len - 1] = 0;
    }
    if (nkinds) *nkinds = n;
    return PSC_OK;
}

int psc_device_bytes(psc_ctx* c, double* bytes) {
    if (!c || !bytes) return PSC_ERR_ARG;
    *bytes = c->dev_bytes;
    return PSC_OK;
}

}  // extern "C"
