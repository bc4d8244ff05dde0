// psc_ctx.hpp -- host-side state of a context (device arrays, launch
// shape, captured graphs, adjoint workspace, kernel timers) shared by the
// translation units of the library: the C ABI (psc_lib.hip) and the launch
// shape instantiations of the stepper (psc_shapes_*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/psc.h"
#include "psc_kernels.hpp"
#include "psc_adjoint.hpp"

namespace psc {


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
    K_Z_FWD_KICK_SYNC,
    K_Z_KICK_GW_SYNC,
    K_PAD_KICK_SYNC,
    K_NKINDS
};
inline const char* const kKindNames[K_NKINDS] = {"z_inverse_drift", "y_inverse", "x_force", "y_forward",
                                    "z_forward_kick", "scalars", "z_inverse", "z_forward_kick_first",
                                    "diagnostics", "z_inverse_drift_gw", "x_force_gw", "z_forward_kick_gw",
                                    "y_inverse_gw", "pad_drift", "pad_z_inverse", "pad_y_inverse", "pad_x_force",
                                    "pad_y_forward", "pad_z_forward", "pad_kick", "z_forward_kick_sync",
                                    "z_forward_kick_gw_sync", "pad_kick_sync"};

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
    virtual int force_chain(psc_ctx* c, bool drift, bool reduce, const KickArgs& ka, bool sync_out) = 0;
    virtual int nV() const = 0;    // partial sums of V per force evaluation
    virtual int nKG() const = 0;   // partial sums of K, G per kick
    virtual size_t padded_elems() const = 0;
    virtual int n_pad() const = 0;   // N_pad, the padded lattice's points per edge
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
    int coupling = 0;                 // PSC_COUPLING_* the captured sweep differentiates
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


}  // namespace psc

using namespace psc;

struct psc_ctx {
    psc_config cfg{};
    int N = 0, B = 1;
    size_t N3 = 0, es = 0;
    size_t MS = 0;   // member stride of the padded device arrays (elements)
    int device = 0;  // the device current at psc_create, which holds the arrays
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
    hipGraphExec_t g_mid = nullptr, g_last = nullptr, g_last_sync = nullptr;
    // synchronised momenta stored by the last step of a call (pi, pi_u pairs,
    // the dealiased stepper's ghost planes) and the context's state: 0 every
    // member synchronised, 1 every member ahead by the look-ahead kick, 2 some
    // may be (a call with the adaptive rule: members at their output time end
    // synchronised)
    void* Ps = nullptr;
    void* PUs = nullptr;
    void* Gps = nullptr;
    int ahead = 0;
    double t_out = 1e300, dt_cap = 1e300;   // targets of the adaptive rule (psc_set_output_times)
    std::unique_ptr<AdjWork> adj;
    double dev_bytes = 0.0;
    int shape = 0;   // index into kShapes
    int pad_shape = -1;   // dealiased stepper: index into kShapes of its padded shape
    int NP = 0;           // dealiased stepper: N_pad (2N or 3N/2), else 0
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

namespace psc {


inline StepConsts consts_of(const psc_ctx* c) {
    StepConsts s{};
    s.dt = c->cfg.dt;
    s.kIR = c->cfg.kIR;
    s.f2m = (c->cfg.fStar / c->cfg.MPl) * (c->cfg.fStar / c->cfg.MPl);
    s.N3 = (double)c->N3;
    s.alpha = c->cfg.alpha;
    s.expansion_static = c->cfg.expansion == PSC_EXPANSION_STATIC;
    s.nflds = c->cfg.nflds;
    s.NV = c->padeng ? (double)c->NP * c->NP * c->NP : (double)c->N3;   // lattice sites V is averaged over
    s.dt_rule = c->cfg.dt_rule;
    s.kmax = c->cfg.kIR * std::sqrt(3.0) * c->N / 2.0;
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

inline int drain_timing(psc_ctx* c) {
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

// Kinds of a recorded step (Engine::record_step).
enum { kStepMid = 0, kStepLastAhead = 1, kStepLastSync = 2 };

struct EngineBase {
    virtual ~EngineBase() = default;
    // nsteps steps, ending one look-ahead kick ahead (psc_step) or synchronised
    virtual int step(psc_ctx* c, int nsteps) = 0;
    virtual int step_synced(psc_ctx* c, int nsteps) = 0;
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
    // member's y rows o0 .. o0 + rows - 1 into the device array D (rows = N:
    // the dense N^3 lattice), laid out and scaled as k_x_density describes.
    virtual int density(psc_ctx* c, int member, int kind, double* D, double scale, int o0, int rows) = 0;
};

// ---- helpers of psc_lib.hip used by the engines ---------------------------

struct Diag {
    double K[2], G[2], Vt[kMaxTerms], mean[2], var[2], sumsq[2], meff[2];
    double V() const {
        double v = 0.0;
        for (int t = 0; t < kMaxTerms; ++t) v += Vt[t];
        return v;
    }
};
int read_state(psc_ctx* c, int member, MemberState* s);
int write_state(psc_ctx* c, int member, const MemberState* s);
int download(psc_ctx* c, const void* src_base, int member, std::vector<double>& full);
int upload(psc_ctx* c, void* dst_base, int member, const std::vector<double>& full);
int diagnose(psc_ctx* c, int member, const MemberState& s, Diag* d);

// Kernel nodes only, with grids, blocks and LDS within the device's limits:
// checked on every captured graph before it is instantiated.
int check_graph(psc_ctx* c, hipGraph_t g, const char* what);

}  // namespace psc
