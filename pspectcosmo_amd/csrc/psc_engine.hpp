// psc_engine.hpp -- the stepper's kernel sequencing for one launch shape:
// Engine<G, GY> runs evolve_VV_k_nopad (evolution.cpp:1131-1198) with the
// GW chain, diagnostics, spectra and the adjoint on the passes of shape G
// (the y passes on GY); PadEngine<GP, GR, N> runs the dealiased stepper's
// force chain (evolve_VV_k_pad, evolution.cpp:1271-1759) of the N lattice on
// the padded shape GP.  The shape instantiations live in psc_shapes_*.hip, one
// group per translation unit, so that they compile in parallel.
#pragma once
#include <algorithm>

#include "psc_ctx.hpp"
#include "psc_padded.hpp"

namespace psc {

// Grid of a multi-tile kernel: as many blocks as the current device keeps
// resident at once (compute units x blocks per unit), at most one per tile.
inline int resident_blocks(const void* kernel, int threads, int tiles) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return std::min(tiles, cus * per_cu);
}

// G drives every pass but the y-axis ones, which run on GY: a plan with more
// columns per block can suit the strided y passes where it would starve the
// x passes (their rows hold fused force, source and reduction work).
template <class G, class GY = G>
struct Engine final : EngineBase {
    using T = typename G::T;
    static constexpr int N = G::N;
    // blocks per member of the force x pass: one per tile, or for the
    // multi-tile form one per resident slot of the device
    const int xblocks = G::XMT ? resident_blocks((const void*)k_x_force_mt<G>, G::THREADS_COL, G::BLOCKS_COL)
                               : G::BLOCKS_COL;
    // the GW chain's merged x pass (k_x_gw) writes one partial per tile
    int nV_blocks(const psc_ctx* c) const { return c->cfg.with_gw ? G::BLOCKS_COL : xblocks; }
    // explicit kick coefficients (the adjoint's replays); prologue: the first
    // kick at a_start only
    static KickArgs kick_args(double c1, double c2, bool prologue = false) {
        KickArgs ka{};
        ka.c1 = c1;
        ka.c2 = c2;
        ka.prologue = prologue ? 1 : 0;
        return ka;
    }
    // the fields of a scalar update that begins substep q of the step
    static void set_begin_weights(ScalarOp& op, int q, const std::vector<double>& w) {
        const int S = (int)w.size();
        op.begin = 1;
        op.w_begin = w[q];
        op.begin_last = q == S - 1;
        op.w_after = w[(q + 1) % S];
    }
    // copy of n complex elements as a kernel (captured graphs hold kernels only)
    static int copy(psc_ctx* c, void* dst, const void* src, size_t n) {
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL((k_copy<T>), dim3(blocks), dim3(256), 0, c->stream, static_cast<cpx<T>*>(dst),
                           static_cast<const cpx<T>*>(src), n);
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }
    static cpx<T>* P(void* p) { return static_cast<cpx<T>*>(p); }

    int launch_z_inverse(psc_ctx* c, bool drift, const cpx<T>* src, cpx<T>* Pf, const cpx<T>* Pp,
                         cpx<T>* W, MemberState* ms, int members) {
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
        return timed(c, K_X_FORCE, 2 * X, [&] {
            if constexpr (G::XMT)
                hipLaunchKernelGGL((k_x_force_mt<G>), dim3(xblocks, 1, n), dim3(G::THREADS_COL), 0, c->stream,
                                   P(c->W) + off, c->ms + m0, P(c->tw), c->partV + (size_t)m0 * xblocks,
                                   P(c->Pf) + off);
            else
                hipLaunchKernelGGL((k_x_force<G>), dim3(N / G::TC, N, n), dim3(G::THREADS_COL), 0, c->stream,
                                   P(c->W) + off, c->ms + m0, P(c->tw), c->partV + (size_t)m0 * G::BLOCKS_COL,
                                   P(c->Pf) + off);
        });
    }
    // z kick of members m0 .. m0 + n - 1; SYNC_OUT also stores the
    // synchronised momenta (the last substep of a call that ends ahead)
    int launch_z_kick(psc_ctx* c, bool reduce, const KickArgs& ka, bool sync_out, int m0 = 0, int n = -1) {
        if (n < 0) n = c->B;
        const size_t off = (size_t)m0 * G::MS;
        double* pkg = c->partKG + (size_t)m0 * G::BLOCKS_ROW * 2;
        const double X = (double)n * G::N3 * sizeof(cpx<T>);
        dim3 grid(G::BLOCKS_ROW, n), block(G::THREADS_ROW);
        const StepConsts sc = consts_of(c);
        const bool st = streams(n);
        auto go = [&](auto red, auto stream, auto sync) {
            hipLaunchKernelGGL((k_z_kick<G, decltype(red)::value, false, decltype(stream)::value, decltype(sync)::value>),
                               grid, block, 0, c->stream, P(c->W) + off, P(c->Pf) + off, P(c->Pp) + off, c->ms + m0,
                               P(c->tw), pkg, ka, sc, nullptr, P(c->Ps) + off);
        };
        using TT = std::true_type;
        using FF = std::false_type;
        const int kind = !reduce ? K_Z_FWD_KICK_NR : (sync_out ? K_Z_FWD_KICK_SYNC : K_Z_FWD_KICK);
        return timed(c, kind, (sync_out ? 5 : 4) * X, [&] {
            if (!reduce) {   // the call's first kick
                if (st) go(FF{}, TT{}, FF{});
                else go(FF{}, FF{}, FF{});
            } else if (sync_out) {
                if (st) go(TT{}, TT{}, TT{});
                else go(TT{}, FF{}, TT{});
            } else {
                if (st) go(TT{}, TT{}, FF{});
                else go(TT{}, FF{}, FF{});
            }
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
    int force_chain_gw(psc_ctx* c, bool drift, bool reduce, const KickArgs& ka, bool sync_out) {
        const size_t astride = (size_t)c->B * G::MS;
        const double X = (double)c->B * G::N3 * sizeof(cpx<T>);
        const StepConsts sc = consts_of(c);
        int rc;
        dim3 rgrid(G::BLOCKS_ROW, c->B), rblock(G::THREADS_ROW);
        // u is drifted by the u kick at the end of the chain (k_z_kick_u)
        if (drift)
            rc = timed(c, K_Z_INV_GW, 5 * X, [&] {
                if (streams(c->B))
                    hipLaunchKernelGGL((k_z_inverse_gw<G, true, false, false, true>), rgrid, rblock, 0, c->stream,
                                       P(c->Pf), P(c->Pp), P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), astride,
                                       c->cfg.kIR);
                else
                    hipLaunchKernelGGL((k_z_inverse_gw<G, true, false, false>), rgrid, rblock, 0, c->stream, P(c->Pf),
                                       P(c->Pp), P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), astride, c->cfg.kIR);
            });
        else
            rc = timed(c, K_Z_INV_GW, 3 * X, [&] {
                hipLaunchKernelGGL((k_z_inverse_gw<G, false>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pp),
                                   P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), astride, c->cfg.kIR);
            });
        if (rc) return rc;
        if ((rc = launch_y_inverse_gw(c, astride))) return rc;
        // the fields' x pass (V') and the gradients' (the source and its x
        // forward transforms) in one pass over the columns
        rc = timed(c, K_X_FORCE_GW, 7 * X, [&] {
            hipLaunchKernelGGL((k_x_gw<G>), dim3(N / G::TC, N, c->B), dim3(G::THREADS_COL), 0, c->stream, P(c->W),
                               c->ms, P(c->tw), c->partV, P(c->Pf), astride, c->cfg.kIR);
        });
        if (rc) return rc;
        if ((rc = launch_y<-1>(c, P(c->W), 4 * c->B))) return rc;
        // the fields' kick (with the reduction) and the u kick, one source
        // array per block
        if ((rc = launch_z_kick(c, reduce, ka, sync_out))) return rc;
        return timed(c, sync_out ? K_Z_KICK_GW_SYNC : K_Z_KICK_GW, (sync_out ? 15 : 12) * X + (drift ? 3 * X : 0), [&] {
            dim3 ugrid(G::BLOCKS_ROW, c->B, 3);
            auto go = [&](auto sync, auto dr, auto st) {
                hipLaunchKernelGGL((k_z_kick_u<G, decltype(sync)::value, decltype(dr)::value, decltype(st)::value>),
                                   ugrid, rblock, 0, c->stream, P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), ka, sc,
                                   astride, decltype(sync)::value ? P(c->PUs) : nullptr);
            };
            using TT = std::true_type;
            using FF = std::false_type;
            auto go_st = [&](auto sync, auto dr) {
                if (streams(c->B)) go(sync, dr, TT{});
                else go(sync, dr, FF{});
            };
            if (sync_out) {
                if (drift) go_st(TT{}, TT{});
                else go_st(TT{}, FF{});
            } else {
                if (drift) go_st(FF{}, TT{});
                else go_st(FF{}, FF{});
            }
        });
    }

    // One force evaluation chain; DRIFT applies the substep's drift first.
    int force_chain(psc_ctx* c, bool drift, bool reduce, const KickArgs& ka, bool sync_out) {
        if (c->padeng) return c->padeng->force_chain(c, drift, reduce, ka, sync_out);
        if (c->cfg.with_gw) return force_chain_gw(c, drift, reduce, ka, sync_out);
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
            if ((rc = launch_z_kick(c, reduce, ka, sync_out, m0, n))) return rc;
        }
        return PSC_OK;
    }

    // All substeps of one step.  kind: kStepMid (an interior step of a call),
    // kStepLastAhead (the last step of a call: its last substep also applies
    // the next step's first kick, the look-ahead, and stores the synchronised
    // momenta), kStepLastSync (the last step of a call that ends synchronised).
    // Kick coefficients come from the member states (set by the scalar update
    // that begins each substep), so the recorded step does not depend on dt.
    int record_step(psc_ctx* c, int kind) {
        const int S = (int)c->w.size();
        for (int p = 0; p < S; ++p) {
            const bool last = p == S - 1;
            KickArgs ka{};
            ka.from_state = 1;
            ka.no_lookahead = (kind == kStepLastSync && last) ? 1 : 0;
            int rc = force_chain(c, true, true, ka, kind == kStepLastAhead && last);
            if (rc) return rc;
            ScalarOp op{};
            op.finish = 1;
            op.last_of_step = last;
            op.begin = (kind == kStepLastSync && last) ? 0 : 1;
            op.w_finish = c->w[p];
            const int q = (p + 1) % S;   // the substep begun next
            op.w_begin = c->w[q];
            op.begin_last = q == S - 1;
            op.w_after = c->w[(q + 1) % S];
            op.end_call = last && kind != kStepMid ? (kind == kStepLastAhead ? 1 : 2) : 0;
            op.nV = c->padeng ? c->padeng->nV() : nV_blocks(c);
            op.nKG = c->padeng ? c->padeng->nKG() : G::BLOCKS_ROW;
            if ((rc = launch_scalars(c, op))) return rc;
        }
        return PSC_OK;
    }

    int build_graph(psc_ctx* c, int kind, hipGraphExec_t* out) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = record_step(c, kind);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return c->fail(PSC_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        if ((rc = check_graph(c, g, "step graph"))) {
            (void)hipGraphDestroy(g);
            return rc;
        }
        HIPCHK(c, hipGraphInstantiate(out, g, nullptr, nullptr, 0));
        HIPCHK(c, hipGraphDestroy(g));
        return PSC_OK;
    }

    // The call's start: members that are not ahead begin substep 0 (dt from
    // the rule) and take its first kick, from a force at the start point.
    // The force is the one the previous call's last substep evaluated at the
    // same fields and a (evolution.cpp:1144-1152), so a call that continues
    // from an ahead state (c->ahead == 1) skips this entirely.
    int prologue(psc_ctx* c) {
        ScalarOp op{};
        op.begin = 1;
        op.prologue = 1;
        const int S = (int)c->w.size();
        op.w_begin = c->w[0];
        op.begin_last = S == 1;
        op.w_after = c->w[1 % S];
        op.nV = c->padeng ? c->padeng->nV() : nV_blocks(c);
        op.nKG = c->padeng ? c->padeng->nKG() : G::BLOCKS_ROW;
        int rc = launch_scalars(c, op);
        if (rc) return rc;
        KickArgs ka{};
        ka.from_state = 1;
        ka.prologue = 1;
        return force_chain(c, false, false, ka, false);
    }

    int step(psc_ctx* c, int nsteps) override { return run_steps(c, nsteps, false); }
    int step_synced(psc_ctx* c, int nsteps) override { return run_steps(c, nsteps, true); }

    // nsteps steps; the call ends ahead (the look-ahead kick applied, the
    // synchronised momenta stored) or, with end_synced, synchronised.
    int run_steps(psc_ctx* c, int nsteps, bool end_synced) {
        if (nsteps <= 0) return PSC_OK;
        int rc;
        if (c->ahead != 1 && (rc = prologue(c))) return rc;
        const int last = end_synced ? kStepLastSync : kStepLastAhead;
        const bool graph = c->cfg.use_graph && !c->timing;
        if (graph) {
            hipGraphExec_t& gl = end_synced ? c->g_last_sync : c->g_last;
            if (nsteps > 1 && !c->g_mid && (rc = build_graph(c, kStepMid, &c->g_mid))) return rc;
            if (!gl && (rc = build_graph(c, last, &gl))) return rc;
            for (int s = 0; s < nsteps - 1; ++s) HIPCHK(c, hipGraphLaunch(c->g_mid, c->stream));
            HIPCHK(c, hipGraphLaunch(gl, c->stream));
        } else {
            for (int s = 0; s < nsteps; ++s)
                if ((rc = record_step(c, s == nsteps - 1 ? last : kStepMid))) return rc;
        }
        HIPCHK(c, hipGetLastError());   // a failed launch is reported here, not by a later call
        c->ahead = end_synced ? 0 : (c->cfg.dt_rule == PSC_DT_FIXED ? 1 : 2);
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
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, c->stream, c->part, G::BLOCKS_COL, kDiagWidth,
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

    int density(psc_ctx* c, int member, int kind, double* D, double scale, int o0, int rows) override {
        if (o0 < 0 || rows < 1 || o0 + rows > N) return c->fail(PSC_ERR_ARG, "bad density row range");
        const size_t off = (size_t)member * G::MS;
        cpx<T>* W = P(c->W) + off;
        const cpx<T>* tw = P(c->tw);
        dim3 cgrid(N / G::TC, rows, 1), cblock(G::THREADS_COL);
        int rc;
        if (kind < 2) {
            const cpx<T>* src = P(kind == 0 ? c->Pf : c->Pp) + off;
            if ((rc = launch_z_inverse(c, false, src, nullptr, nullptr, W, c->ms + member, 1))) return rc;
            if ((rc = launch_y<+1>(c, W, 1))) return rc;
            if (kind == 0)
                hipLaunchKernelGGL((k_x_density<G, 0>), cgrid, cblock, 0, c->stream, W, D, c->ms + member, tw,
                                   src, scale, o0, rows);
            else
                hipLaunchKernelGGL((k_x_density<G, 1>), cgrid, cblock, 0, c->stream, W, D, c->ms + member, tw,
                                   src, scale, o0, rows);
        } else {
            for (int d = 0; d < 3; ++d) {
                hipLaunchKernelGGL((k_z_inverse_deriv<G>), dim3(G::BLOCKS_ROW, 1), dim3(G::THREADS_ROW), 0,
                                   c->stream, P(c->Pf) + off, W, tw, d, c->cfg.kIR);
                if ((rc = launch_y<+1>(c, W, 1))) return rc;
                if (d == 0)
                    hipLaunchKernelGGL((k_x_density<G, 3>), cgrid, cblock, 0, c->stream, W, D, c->ms + member, tw,
                                       nullptr, scale, o0, rows);
                else
                    hipLaunchKernelGGL((k_x_density<G, 2>), cgrid, cblock, 0, c->stream, W, D, c->ms + member, tw,
                                       nullptr, scale, o0, rows);
            }
        }
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    // ---- adjoint helpers --------------------------------------------------
    static constexpr int kObjBlocks = (int)(((size_t)N * N * (N / 2 + 1) + 255) / 256);
    int adj_workspace(psc_ctx* c, int K, int nseg, bool gw, int coupling, AdjWork** out) {
        const int S = (int)c->w.size();
        AdjWork* cur = c->adj.get();
        if (cur && cur->K == K && cur->S == S && cur->nseg >= nseg && cur->gw == gw && cur->coupling == coupling) {
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
        w->coupling = coupling;
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
        int rc = copy(c, w.p[0], c->Pp, G::MS);
        if (rc) return rc;
        ScalarOp op{};
        set_begin_weights(op, 0, c->w);
        op.prologue = 1;
        op.nV = G::BLOCKS_COL;
        op.nKG = G::BLOCKS_ROW;
        if ((rc = launch_scalars(c, op, w.rms))) return rc;
        dim3 rgrid(G::BLOCKS_ROW, 1), rblock(G::THREADS_ROW), cgrid(N / G::TC, N, 1), cblock(G::THREADS_COL);
        // prologue force at (f_0, a_0): first kick, real-space fields of f_0
        hipLaunchKernelGGL((k_z_inverse<G, false, false, true>), rgrid, rblock, 0, c->stream, P(c->Pf), P(c->Pf),
                           P(c->Pp), P(c->W), c->ms, P(c->tw), nullptr, nullptr);
        if ((rc = launch_y<+1>(c, P(c->W), 1))) return rc;
        hipLaunchKernelGGL((k_x_force<G, true>), cgrid, cblock, 0, c->stream, P(c->W), c->ms, P(c->tw), c->partV,
                           P(c->Pf), P(w.r[0]));
        if ((rc = launch_y<-1>(c, P(c->W), 1))) return rc;
        hipLaunchKernelGGL((k_z_kick<G, false>), rgrid, rblock, 0, c->stream, P(c->W), P(c->Pf), P(c->Pp), c->ms,
                           P(c->tw), c->partKG, kick_args(c->w[0] * dt / 2.0, 0.0, true), sc, nullptr, nullptr);
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
                               P(c->Pp), c->ms, P(c->tw), c->partKG,
                               kick_args(c->w[p] * dt / 2.0, last ? 0.0 : wn * dt / 2.0), sc, P(w.p[j + 1]), nullptr);
            ScalarOp o{};
            o.finish = 1;
            o.last_of_step = p == S - 1;
            if (!last) set_begin_weights(o, (p + 1) % S, c->w);
            o.w_finish = c->w[p];
            o.nV = G::BLOCKS_COL;
            o.nKG = G::BLOCKS_ROW;
            if ((rc = launch_scalars(c, o, w.rms + j + 1))) return rc;
        }
        if ((rc = copy(c, w.f[steps * S], c->Pf, G::MS))) return rc;
        HIPCHK(c, hipGetLastError());
        return PSC_OK;
    }

    // The same replay through the GW chain (force_chain_gw), recording also u
    // and the real-space gradients at every boundary and pi_u half per substep.
    int record_segment_gw(psc_ctx* c, int steps, AdjWork& w) {
        const int S = (int)c->w.size();
        const double dt = c->cfg.dt;
        const StepConsts sc = consts_of(c);
        const size_t as = G::MS;   // one member: arrays MS apart
        const double kIR = c->cfg.kIR;
        int rc = copy(c, w.p[0], c->Pp, as);
        if (rc || (rc = copy(c, w.u[0], c->U, 3 * as))) return rc;
        ScalarOp op{};
        set_begin_weights(op, 0, c->w);
        op.prologue = 1;
        op.nV = G::BLOCKS_COL;
        op.nKG = G::BLOCKS_ROW;
        if ((rc = launch_scalars(c, op, w.rms))) return rc;
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
                           P(c->U), P(c->PU), c->ms, P(c->tw), c->partKG, kick_args(c->w[0] * dt / 2.0, 0.0, true), sc,
                           as, nullptr, nullptr, nullptr);
        for (int j = 0; j < steps * S; ++j) {
            const int p = j % S;
            const bool last = j == steps * S - 1;
            const double wn = c->w[(p + 1) % S];
            // the u drift and its recording (u after the drift, pi_u half) run in
            // the kick, which loads u and pi_u anyway, as in the forward chain
            hipLaunchKernelGGL((k_z_inverse_gw<G, true, true, false>), rgrid, rblock, 0, c->stream, P(c->Pf),
                               P(c->Pp), P(c->W), P(c->U), P(c->PU), c->ms, P(c->tw), as, kIR, P(w.f[j]),
                               P(w.ph[j]), nullptr, nullptr);
            if ((rc = middle(j + 1))) return rc;
            hipLaunchKernelGGL((k_z_kick_gw<G, true, true, false, true>), rgrid, rblock, 0, c->stream, P(c->W),
                               P(c->Pf), P(c->Pp), P(c->U), P(c->PU), c->ms, P(c->tw), c->partKG,
                               kick_args(c->w[p] * dt / 2.0, last ? 0.0 : wn * dt / 2.0), sc, as, P(w.p[j + 1]),
                               nullptr, nullptr, P(w.u[j + 1]), P(w.puh[j]));
            ScalarOp o{};
            o.finish = 1;
            o.last_of_step = p == S - 1;
            if (!last) set_begin_weights(o, (p + 1) % S, c->w);
            o.w_finish = c->w[p];
            o.nV = G::BLOCKS_COL;
            o.nKG = G::BLOCKS_ROW;
            if ((rc = launch_scalars(c, o, w.rms + j + 1))) return rc;
        }
        if ((rc = copy(c, w.f[steps * S], c->Pf, as))) return rc;
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
                           w.p3, w.coupling);
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
    // full segments of the <chi^2> adjoint go through a graph captured on
    // first use when steps are graph-captured.  The GW objective's segments
    // are launched eagerly: a captured GW segment once ended in a
    // malformed-packet abort whose cause in this library was not found, and
    // eager launches measured no slower there (backward 240 ms eager, 239 ms
    // captured, 128^3 FP32).
    int run_segment(psc_ctx* c, int steps, AdjWork& w, bool bwd) {
        if (steps != w.K || !c->cfg.use_graph || w.gw)
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
            if ((rc = check_graph(c, g, bwd ? "adjoint sweep graph" : "adjoint recording graph"))) {
                (void)hipGraphDestroy(g);
                return rc;
            }
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

// The dealiased stepper of the N lattice on the padded shape GP (N_pad =
// GP::N: 2N or 3N/2); the two z-axis row passes (padding, extraction) may run
// on another plan GR.
template <class GP, class GR, int N, class GPY = GP>
struct PadEngine final : PadBase {
    using T = typename GP::T;
    static constexpr int NP = GP::N;
    static constexpr int H = N / 2;   // the kept modes |n_d| <= H
    static_assert(GR::N == NP && std::is_same_v<typename GR::T, T>, "row plan of another lattice");
    static_assert(GPY::N == NP && std::is_same_v<typename GPY::T, T>, "y plan of another lattice");
    static_assert(pad_pair_ok<GP, N>() && pad_pair_ok<GR, N>(), "padded shape and lattice do not pair");
    static cpx<T>* P(void* p) { return static_cast<cpx<T>*>(p); }
    static constexpr double XN = (double)N * N * N * sizeof(cpx<T>);
    static constexpr double XP = (double)NP * NP * NP * sizeof(cpx<T>);
    static constexpr double KP = (N + 1.0) / NP;   // share of a padded axis holding modes
    // blocks of the padded x pass (multi-tile form: one per resident slot)
    const int xblocks = GP::XMT ? resident_blocks((const void*)k_x_force_mt<GP, true, H>, GP::THREADS_COL,
                                                  GP::BLOCKS_COL)
                                : GP::BLOCKS_COL;

    int nV() const override { return xblocks; }
    int n_pad() const override { return NP; }
    // the padded array is read once after each write and exceeds the
    // memory-side cache: its z passes use non-temporal accesses
    static constexpr bool stream_padded() { return (double)GR::MS * sizeof(cpx<T>) > 96.0 * (1 << 20); }
    // the y passes too, where a tile row covers a whole 128-byte line (with
    // 64-byte rows the two halves of a line come from neighbouring blocks and
    // only ordinary accesses let the cache join them: 8-column 768-point y
    // passes 1214 / 1275 -> 1460 / 1601 us non-temporal)
    static constexpr bool stream_padded_y() { return stream_padded() && GPY::TC * sizeof(cpx<T>) >= 128; }
    // the extracted V' residual, (N + 1)^3 elements, read once by the kick
    static constexpr bool stream_residual() { return (double)(N + 1) * (N + 1) * (N + 1) * sizeof(cpx<T>) > 96.0 * (1 << 20); }
    int nKG() const override { return (int)((size_t)N * N * N / 2048); }
    size_t padded_elems() const override { return GP::MS; }

    // padding with the z-axis inverse, then the y-axis inverse, then V' and
    // the sum of V in the x pass
    int front(psc_ctx* c) {
        int rc = timed(c, K_PAD_Z_INV, XN + XP * KP * KP, [&] {
            if constexpr (stream_padded())
                hipLaunchKernelGGL((k_pad_z<GR, N, true>), dim3(pad_z_blocks<GR, N>()), dim3(GR::THREADS_ROW), 0,
                                   c->stream, P(c->Pf), P(c->Gf), P(c->Wp), P(c->dcp), P(c->twp));
            else
                hipLaunchKernelGGL((k_pad_z<GR, N>), dim3(pad_z_blocks<GR, N>()), dim3(GR::THREADS_ROW), 0, c->stream,
                                   P(c->Pf), P(c->Gf), P(c->Wp), P(c->dcp), P(c->twp));
        });
        if (rc) return rc;
        // the y passes and the stored x rows cover the padded x planes |x| <= N/2;
        // the y inverse reads the rows |y| <= N/2 of them
        rc = timed(c, K_PAD_Y_INV, XP * KP * (KP + 1.0), [&] {
            hipLaunchKernelGGL((k_strided<GPY, 1, +1, true, false, H, stream_padded_y()>), dim3(NP / GPY::TC, N + 1, 1),
                               dim3(GPY::THREADS_COL), 0, c->stream, P(c->Wp), P(c->twp));
        });
        if (rc) return rc;
        return timed(c, K_PAD_X_FORCE, XP * KP * 2.0, [&] {
            if constexpr (GP::XMT)
                hipLaunchKernelGGL((k_x_force_mt<GP, true, H>), dim3(xblocks, 1, 1), dim3(GP::THREADS_COL), 0,
                                   c->stream, P(c->Wp), c->ms, P(c->twp), c->partV, P(c->dcp));
            else
                hipLaunchKernelGGL((k_x_force<GP, false, true, H>), dim3(NP / GP::TC, NP, 1), dim3(GP::THREADS_COL),
                                   0, c->stream, P(c->Wp), c->ms, P(c->twp), c->partV, P(c->dcp));
        });
    }

    int force_chain(psc_ctx* c, bool drift, bool reduce, const KickArgs& ka, bool sync_out) override {
        int rc;
        if (drift && (rc = timed(c, K_PAD_DRIFT, 3 * XN, [&] {
                hipLaunchKernelGGL((k_drift_pad<T, N>), dim3((unsigned)((size_t)N * N * N / 256)), dim3(256), 0,
                                   c->stream, P(c->Pf), P(c->Pp), P(c->Gf), P(c->Gp), c->ms);
            })))
            return rc;
        if ((rc = front(c))) return rc;
        rc = timed(c, K_PAD_Y_FWD, XP * KP * (1.0 + KP), [&] {
            hipLaunchKernelGGL((k_strided<GPY, 1, -1, true, false, H, stream_padded_y()>), dim3(NP / GPY::TC, N + 1, 1),
                               dim3(GPY::THREADS_COL), 0, c->stream, P(c->Wp), P(c->twp));
        });
        if (rc) return rc;
        rc = timed(c, K_PAD_Z_FWD, XN * (N + 1) * (N + 1) / ((double)N * N) * 3, [&] {
            if constexpr (stream_padded())
                hipLaunchKernelGGL((k_unpad_z<GR, N, true>), dim3(pad_z_blocks<GR, N>()), dim3(GR::THREADS_ROW), 0,
                                   c->stream, P(c->Wp), P(c->Rp), P(c->twp));
            else
                hipLaunchKernelGGL((k_unpad_z<GR, N>), dim3(pad_z_blocks<GR, N>()), dim3(GR::THREADS_ROW), 0,
                                   c->stream, P(c->Wp), P(c->Rp), P(c->twp));
        });
        if (rc) return rc;
        const StepConsts sc = consts_of(c);
        return timed(c, sync_out ? K_PAD_KICK_SYNC : K_PAD_KICK, (sync_out ? 5 : 4) * XN, [&] {
            constexpr bool st = stream_residual();
            if (reduce && sync_out)
                hipLaunchKernelGGL((k_kick_pad<GP, N, true, true, st>), dim3(nKG()), dim3(256), 0, c->stream, P(c->Rp),
                                   P(c->Pf), P(c->Pp), P(c->Gf), P(c->Gp), c->ms, c->partKG, ka, sc, P(c->Ps),
                                   P(c->Gps));
            else if (reduce)
                hipLaunchKernelGGL((k_kick_pad<GP, N, true, false, st>), dim3(nKG()), dim3(256), 0, c->stream,
                                   P(c->Rp), P(c->Pf), P(c->Pp), P(c->Gf), P(c->Gp), c->ms, c->partKG, ka, sc, nullptr,
                                   nullptr);
            else
                hipLaunchKernelGGL((k_kick_pad<GP, N, false, false, st>), dim3(nKG()), dim3(256), 0, c->stream,
                                   P(c->Rp), P(c->Pf), P(c->Pp), P(c->Gf), P(c->Gp), c->ms, c->partKG, ka, sc, nullptr,
                                   nullptr);
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
        *V = v / ((double)NP * NP * NP);
        return PSC_OK;
    }
};

// Launch shapes: index 0 is the default of each (precision, N); further
// entries are the candidates the autotuner times (same results up to the
// rounding of a different radix plan; identical when only TC / L differ).

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
    int rc = adj_workspace(c, K, nseg, gw, sp->coupling, &w);
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
        c->ahead = 0;   // checkpoints hold synchronised states
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
        if ((rc = sg == nseg - 1 ? run_segment(c, steps, *w, false) : step_synced(c, steps))) return rc;
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
        // <dV/d(coupling)> from the recorded final real-space fields
        const int nsub_last = (M - (nseg - 1) * K) * S;
        hipLaunchKernelGGL((k_vg_sums<G>), dim3(256), dim3(256), 0, c->stream, P(w->r[nsub_last]), c->ms,
                           sp->coupling, c->part);
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
    const Couplings<double> cp0(s0.par);
    const double f0[2] = {sp->phi0, sp->chi0};
    const double V0 = potential_sum<Model, double>(c->cfg.nflds, f0, cp0);
    const double H = std::sqrt((sp->dphi0 * sp->dphi0 / 2.0 + V0) * sc.f2m / 3.0);
    const double vphi = Model::dvdf<double>(0, f0, cp0);
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
    // the coupling also enters da/dt(0) = a0 H(phi0): dH/dc = f2m dV/dc / (6 H)
    out[1] = hst.g2bar + (sc.expansion_static ? 0.0
                                               : hst.bbar * s0.a * sc.f2m *
                                                     dV_dcoupling<Model, double>(sp->coupling, f0, cp0) / (6.0 * H));
    out[2] = phi0bar;
    out[3] = fwd_ms;
    out[4] = bwd_ms;
    out[5] = w->bytes;
    return PSC_OK;
}


template <typename T, int N, class R, int TC, int L>
using S = Shape<T, N, R, TC, L>;

struct ShapeEntry {
    int precision, N;
    const char* name;
    EngineBase* (*make)();    // the stepper of the N lattice, or null (padded stepper only)
    PadBase* (*make_pad)(int n);   // the dealiased stepper of the n lattice on this padded
                                   // shape (n = N/2, or 2N/3 when 3 divides N), or null
};

template <class G, class GY = G>
EngineBase* make_one() { return new Engine<G, GY>(); }

// N_pad = 2N (the reference's shipped pair, parameters.h:22-23) and, for the
// lengths 3 * 2^k, N_pad = 3N/2 (the two-thirds rule of the quadratic terms).
constexpr bool lattice_length(int n) {   // 2^k (16 .. 512) or 3 * 2^k (48 .. 384) of an unpadded lattice
    return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 48 || n == 96 ||
           n == 192 || n == 384;
}

template <class G, class GR = G, class GY = G>
PadBase* make_pad_one(int n) {
    constexpr int NP = G::N;
    if constexpr (lattice_length(NP / 2) && NP % 2 == 0 && pad_pair_ok<G, NP / 2>() && pad_pair_ok<GR, NP / 2>())
        if (2 * n == NP) return new PadEngine<G, GR, NP / 2, GY>();
    if constexpr (NP % 3 == 0 && lattice_length(2 * NP / 3) && pad_pair_ok<G, 2 * NP / 3>() &&
                  pad_pair_ok<GR, 2 * NP / 3>())
        if (3 * n == 2 * NP) return new PadEngine<G, GR, 2 * NP / 3, GY>();
    (void)n;
    return nullptr;
}

#define PSC_SHAPE(P, T, N, NAME, ...) {P, N, NAME, &make_one<S<T, N, __VA_ARGS__>>, &make_pad_one<S<T, N, __VA_ARGS__>>}
// A shape whose y passes run on the second plan GY; not offered to the
// dealiased stepper, whose y passes are its own.
#define PSC_SHAPE_Y(P, NAME, G, GY) {P, G::N, NAME, &make_one<G, GY>, nullptr}
// A shape of the padded stepper only, its z row passes on GR.
#define PSC_SHAPE_PAD(P, NAME, G, GR) {P, G::N, NAME, nullptr, &make_pad_one<G, GR>}
// ... and its y passes on GY.
#define PSC_SHAPE_PAD_Y(P, NAME, G, GR, GY) {P, G::N, NAME, nullptr, &make_pad_one<G, GR, GY>}
// A shape of the unpadded stepper only.
#define PSC_SHAPE_ONLY(P, T, N, NAME, ...) {P, N, NAME, &make_one<S<T, N, __VA_ARGS__>>, nullptr}

}  // namespace psc
