/*
 * psc.h -- C ABI of the MI355X-native PSpectCosmo lattice time step.
 *
 * The reference (PSpectCosmo) has no plugin or FFI: its boundary is the
 * compile-time configuration (parameters.h, model.h), the step function
 * `void evolve_VV_k_nopad(double dt)` acting on global state
 * (pspectcosmo.h:241-242, evolution.cpp:1131-1198), and the output routines
 * that read the same globals (output.cpp).  Every entry point below states the
 * reference routine or data it replaces.
 *
 * Conventions (shared by every function):
 *   - plain C types only; every function returns an int status (PSC_OK == 0)
 *     and never exits the process (the reference prints and calls exit(1),
 *     output.cpp:450-454, 480-483);
 *   - k-space arrays crossing the boundary are FP64 half spectra in FFTW r2c
 *     layout: element [(i*N + j)*(N/2+1) + k] holds (re, im) of the mode with
 *     lattice momentum (i<=N/2 ? i : i-N, j<=N/2 ? j : j-N, k), unnormalised
 *     (forward e^{-ikx}), exactly as f_k / pi_k in pspectcosmo.h:73-74;
 *   - real-space arrays are FP64, index [(i*N + j)*N + k] (pspectcosmo.h:63);
 *   - one context per host thread; device memory is owned by the context;
 *     calls on one context are not re-entrant.
 */
#ifndef PSC_H
#define PSC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSC_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
#define PSC_OK 0
#define PSC_ERR_ARG 1          /* invalid argument or configuration            */
#define PSC_ERR_DEVICE 2       /* HIP runtime error or missing GPU              */
#define PSC_ERR_NOMEM 3        /* device or host allocation failed             */
#define PSC_ERR_STATE 4        /* call out of order (e.g. step before load)    */
#define PSC_ERR_NONFINITE 5    /* field mean not finite: output.cpp:450-454    */
#define PSC_ERR_NEGATIVE_A 6   /* scale factor < 0: output.cpp:480-483         */
#define PSC_ERR_UNSUPPORTED 7  /* configuration outside what the build covers */

/* ---- enumerations ------------------------------------------------------ */
#define PSC_VV2 2              /* velocity Verlet, pspectcosmo.h:283-285       */
#define PSC_Y4 4               /* Yoshida 4th order, pspectcosmo.h:286-288     */

#define PSC_FP64 0             /* all fields/FFTs in FP64                      */
#define PSC_FP32 1             /* FP32 fields/FFTs, FP64 scale factor,
                                  reductions and spectrum bins               */

#define PSC_EXPANSION_STATIC 0 /* a = 1, da/dt = 0 (Minkowski test mode)       */
#define PSC_EXPANSION_SELF 2   /* self-consistent Friedmann, parameters.h:65   */

#define PSC_DT_FIXED 0         /* psc_config.dt_rule                           */
#define PSC_DT_REFERENCE 1

#define PSC_GW_PROJ_REFERENCE 0 /* one-sided P.u as in output.cpp:819-831      */
#define PSC_GW_PROJ_TT 1        /* Lambda_ij,lm u_lm, true TT projection       */

/* Number of doubles written by psc_energies / psc_oracle energies. */
#define PSC_ENERGY_LEN 16
/* Columns of psc_spectra per bin: p, S_f, S_pi, n_k, numpoints. */
#define PSC_SPECTRA_COLS 5
/* Columns of psc_gw_spectrum per bin: p, S_h, Omega_gw, prefactor, numpoints. */
#define PSC_GW_SPECTRA_COLS 5

/* Run configuration: the runtime form of parameters.h (parameters.h:10-134). */
typedef struct psc_config {
    int N;             /* lattice points per edge, 2^k or 3*2^k in 16..512 (parameters.h:22) */
    int nflds;         /* 1 or 2 scalar fields (parameters.h:20)                          */
    int integrator;    /* PSC_VV2 or PSC_Y4 (parameters.h:10 `VV`)                        */
    int precision;     /* PSC_FP64 or PSC_FP32                                             */
    int batch;         /* independent ensemble members advanced together (>= 1)           */
    int with_gw;       /* evolve the 6 symmetric u_ij and pi_u_ij (parameters.h:15)       */
    int alpha;         /* time variable exponent; 0 = cosmic time (parameters.h:121)      */
    int expansion;     /* PSC_EXPANSION_SELF (reference) or PSC_EXPANSION_STATIC          */
    int yoshida_exact; /* 0: the reference's Yoshida digits (pspectcosmo.h:288);
                          1: the exact triple-jump weights                               */
    int use_graph;     /* capture each step in a hipGraph and replay it                   */
    int autotune;      /* 1: adopt the launch shape cached by psc_autotune for this
                          device / N / precision / batch / GW, if any                   */
    int pad;           /* 0: evolve_VV_k_nopad; 2: the dealiased stepper evolve_VV_k_pad
                          (PAD_FLAG, evolution.cpp:1271-1759) on the N_pad lattice     */
    int dt_rule;       /* PSC_DT_FIXED: every step takes dt; PSC_DT_REFERENCE: each step
                          takes the reference's calculate_dt, capped and clamped to the
                          output time of psc_set_output_times (pspectcosmo.cpp:213-233,
                          622-627), computed on the device                           */
    int n_pad;         /* N_pad of the dealiased stepper (parameters.h:23), larger than N
                          (pspectcosmo.cpp:244): 2N or 3N/2 (N_pad <= 768); 0 = 2N;
                          must be 0 with pad = 0                                      */
    int reserved_i[2];
    double kIR;        /* lattice IR wavenumber, L = 2 pi / kIR (parameters.h:28,32)       */
    double dt;         /* time step in program units (parameters.h:36)                     */
    double fStar;      /* field rescaling (parameters.h:133)                               */
    double omegaStar;  /* time rescaling (parameters.h:134)                                */
    double MPl;        /* reduced Planck mass in program units (parameters.h:122)         */
    double reserved_d[3];
} psc_config;

/* Couplings of the potential, per ensemble member.  The potential itself is
 * compiled in from a model header, as the reference compiles model.h
 * (model.h:34-104): each model header states which couplings it reads.  The
 * built-in models are
 *   preheat      (libpsc.so)             V = m2 phi^2 / 2 + g2 phi^2 chi^2 / 2;
 *                with one field the reference's shipped phi^2 model.h;
 *   lambda_phi4  (libpsc_lambda_phi4.so) V = lambda phi^4 / 4 + g2 phi^2 chi^2 / 2. */
typedef struct psc_model {
    double m2;         /* mass term                                            */
    double g2;         /* phi^2 chi^2 coupling                                 */
    double lambda;     /* phi^4 self-coupling                                  */
    double c[5];       /* further couplings of a user model                    */
} psc_model;

/* Coupling indices (psc_adjoint_spec.coupling): the derivative of the
 * adjoint's objective is taken with respect to this member of psc_model. */
#define PSC_COUPLING_G2 0
#define PSC_COUPLING_M2 1
#define PSC_COUPLING_LAMBDA 2
#define PSC_COUPLING_C0 3      /* c[i] is PSC_COUPLING_C0 + i                  */

/* Initial-condition parameters (parameters.h:62,124-125; initialize.cpp:624-668). */
typedef struct psc_ic_params {
    double f0[2];      /* homogeneous field values, program units (initfield/fStar)        */
    double fd0[2];     /* homogeneous field velocities, program units                      */
    double a0;         /* initial scale factor (parameters.h:29)                           */
    int seed;          /* Park-Miller seed (parameters.h:62); < 1 is the reference's
                          debug mode where every draw returns 0.33 (initialize.cpp:15) */
    int reserved_i;
    double kcutoff;    /* low-pass of the initial modes (parameters.h:31): modes with
                          k^2 > kcutoff^2 are zero and take no draws, so the later
                          modes take the draws they would have used
                          (initialize.cpp:138-145, 281-294); 0, or a value of at
                          least N kIR, cuts nothing                               */
    double reserved_d[3];
} psc_ic_params;

/* ---- host-side utilities (no device needed) --------------------------- */

int psc_abi_version(void);
const char* psc_status_string(int status);
/* Name and formula of the model the library was built with
 * (modelinfo_, model.h:10-19). */
const char* psc_model_name(void);
const char* psc_model_description(void);
/* Number of potential terms of the model with nflds fields
 * (num_potential_terms, model.h:34). */
int psc_num_potential_terms(int nflds);

/* Fill the initial half spectra for all fields, replacing initialize()
 * (initialize.cpp:562-810, non-inflation Rayleigh branch initialize.cpp:181-197).
 * f_k and pi_k hold nflds * N*N*(N/2+1) complex values (2 doubles each).
 * scalars_out[0..3] = a, da/dt, the printed d2a/dt2 (initialize.cpp:667), H. */
int psc_generate_ic(const psc_config* cfg, const psc_model* model, const psc_ic_params* ic,
                    double* f_k, double* pi_k, double* scalars_out);

/* Fill the 6 symmetric GW components (11,22,33,12,13,23) of u_ij and pi_u_ij,
 * replacing initialize_GW() (initialize.cpp:341-470).  The Park-Miller stream
 * continues after the scalar-field draws of psc_generate_ic with the same seed,
 * as in the reference (pspectcosmo.cpp:458,488).  Each output holds
 * 6 * N*N*(N/2+1) complex values. */
int psc_generate_gw_ic(const psc_config* cfg, const psc_model* model, const psc_ic_params* ic,
                       double* u_k, double* pi_u_k);

/* ---- device context ----------------------------------------------------- */

typedef struct psc_ctx psc_ctx;

/* Allocate device state for cfg->batch members (pspectcosmo.cpp:253-282). */
int psc_create(const psc_config* cfg, psc_ctx** out);
int psc_destroy(psc_ctx* ctx);
/* Text of the last error on this context (never NULL). */
const char* psc_last_error(const psc_ctx* ctx);

/* Per-member couplings (model.h:35-98). */
int psc_set_model(psc_ctx* ctx, int member, const psc_model* model);

/* Copy initial half spectra to the device and compute d2a/dt2 from them as
 * initialize.cpp:797 does with kernel_a_k_nopad (evolution.cpp:1061-1090).
 * In an FP32 context the homogeneous mode (k = 0) keeps its FP64 value: the
 * FP32 array holds its rounding and the context the remainder. */
int psc_load_ic_k(psc_ctx* ctx, int member, const double* f_k, const double* pi_k,
                  double a, double ad);
/* Initial conditions generated on the device straight into member `member`:
 * the modes of psc_generate_ic (initialize(), initialize.cpp:562-810, in its
 * visiting order with the same Park-Miller draws; the model set with
 * psc_set_model), then the state as psc_load_ic_k leaves it.  Values agree
 * with psc_generate_ic to within a few units in the last place (the device's
 * log, sin and cos).  scalars_out (may be NULL) as psc_generate_ic. */
int psc_generate_ic_device(psc_ctx* ctx, int member, const psc_ic_params* ic, double* scalars_out);
/* GW modes on the device, as psc_generate_gw_ic followed by psc_load_gw_k
 * (initialize_GW, initialize.cpp:341-470). */
int psc_generate_gw_ic_device(psc_ctx* ctx, int member, const psc_ic_params* ic);
/* Load a member from real-space fields and momenta (nflds x N^3 each, the
 * layout of pspectcosmo.h:63) and t, a, da/dt, as the continue_run branch
 * of the reference reads its lattice image (initialize.cpp:590-616); the
 * spectra are formed on the device and d2a/dt2 is recomputed. */
int psc_load_fields(psc_ctx* ctx, int member, const double* f, const double* pi, double t, double a,
                    double ad);
/* Overwrite the background scalars t, a, da/dt, d2a/dt2 of one member. */
int psc_set_scalars(psc_ctx* ctx, int member, double t, double a, double ad, double add);
/* out[0..3] = t, a, da/dt, d2a/dt2 (pspectcosmo.h:44-45).  Reads the state
 * without synchronising the momenta (see psc_step). */
int psc_get_scalars(psc_ctx* ctx, int member, double* out);
/* Steps of nonzero dt a member has taken since its data were loaded (the
 * driver's step counter, pspectcosmo.cpp:683). */
int psc_get_step_count(psc_ctx* ctx, int member, long long* steps);

/* Advance every member by nsteps steps; replaces `evolve_VV_k_nopad(dt)`
 * called from the main loop (pspectcosmo.cpp:636).  With PSC_DT_FIXED each
 * step takes dt; with PSC_DT_REFERENCE each step's dt is the reference's rule
 * evaluated on the device, and a member that has reached its output time
 * takes steps of zero for the rest of the call.  The call leaves the momenta
 * one look-ahead kick ahead (the first kick of the next step, which the
 * reference recomputes from the same fields, evolution.cpp:1144-1152), so
 * that a following psc_step continues without a force evaluation; every
 * function that reads or changes the state first restores the synchronised
 * momenta, which the last step also stored.  psc_step(n) and n calls of
 * psc_step(1) give bit-identical states. */
int psc_step(psc_ctx* ctx, int nsteps);
/* Change the fixed time step (PSC_DT_FIXED) for later psc_step calls. */
int psc_set_dt(psc_ctx* ctx, double dt);
/* PSC_DT_REFERENCE: the next output time t_out that steps land on (the
 * earlier of the reference's next_Output_Freq_time and
 * next_Output_Infreq_time, pspectcosmo.cpp:624-627) and the largest dt
 * (tOutputFreq), for every member.  A huge t_out leaves the rule unclamped. */
int psc_set_output_times(psc_ctx* ctx, double t_out, double dt_cap);
/* Block until all device work of the context is done. */
int psc_sync(psc_ctx* ctx);

/* Read the state back as FP64 half spectra (f_k, pi_k of pspectcosmo.h:73-74);
 * in an FP32 context the homogeneous mode includes its FP64 remainder. */
int psc_get_k(psc_ctx* ctx, int member, double* f_k, double* pi_k);
/* Real-space fields f and pi (c2r and /N^3, output.cpp:401-425). Either may be NULL. */
int psc_get_fields(psc_ctx* ctx, int member, double* f, double* pi);

/* Energy densities on the lattice as the snapshots write them, each N^3
 * doubles x-major (the reference's [i][j][k]); any pointer may be NULL.
 * pot: sum of the potential terms per site; kin: a^6/2 sum_f pi_f^2
 * (kinetic_energy_lattice, evolution.cpp:7-19); grad: a^2/2 sum_f sum_d
 * (d_d f)^2 with spectral first derivatives whose Nyquist component is zeroed
 * (gradient_energy_lattice, evolution.cpp:23-138); tot = pot + kin + grad
 * (snapshots, output.cpp:1614-1625).  The a-powers are the reference's. */
int psc_energy_densities(psc_ctx* ctx, int member, double* pot, double* kin, double* grad, double* tot);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): the same densities,
 * bit for bit, staged through a fixed 32 MiB device array of the library (one
 * per GPU, held until the process ends) in blocks of y rows instead of a
 * per-call N^3 allocation; it calls no hipMalloc or hipFree.  tot needs pot,
 * kin and grad (there is no host staging either).  Calls that overlap on one
 * GPU share the array and race; PSC_ERR_UNSUPPORTED when N^2 doubles exceed
 * it (N > 2048). */
int psc_energy_densities_noalloc(psc_ctx* ctx, int member, double* pot, double* kin, double* grad, double* tot);

/* Energies, replacing energy_density()/energy() (output.cpp:57-115,1069-1187).
 * out[PSC_ENERGY_LEN]: K_0, G_0, K_1, G_1, V_term0, V_term1 (the first two
 * potential terms; psc_potential_terms has all of them), rho_total,
 * constraint violation |H^2 - rho/3| / |H^2 + rho/3| (output.cpp:1180-1184),
 * a, da/dt, d2a/dt2, t, then <f_0>, <f_1>, var f_0, var f_1 (output.cpp:380-461). */
int psc_energies(psc_ctx* ctx, int member, double* out);

/* Lattice averages of every potential term (potential_energy(term, NULL),
 * model.h:38-52): out[t] for t < psc_num_potential_terms(nflds). */
int psc_potential_terms(psc_ctx* ctx, int member, double* out);

/* Means and variances as written by meansvars() (output.cpp:380-461), per
 * field q: out[6q .. 6q+5] = <f>, a^{alpha-3} <pi>, <f^2>, a^{2(alpha-3)} <pi^2>,
 * sigma_f, a^{alpha-3} sigma_pi (<pi>, <pi^2> from the k = 0 mode and Parseval). */
int psc_means(psc_ctx* ctx, int member, double* out);

/* Binned scalar spectra of field `fld`, replacing spectra() (output.cpp:490-705).
 * out holds nbins rows of PSC_SPECTRA_COLS doubles for bins 1..nbins, where
 * nbins = maxnumbins - 1 and maxnumbins = floor(sqrt(3) * N / 2) + 2
 * (parameters.h:25). */
int psc_spectra(psc_ctx* ctx, int member, int fld, double* out, int nbins);

/* GW state in and out (6 symmetric components, FFTW layout, FP64). */
int psc_load_gw_k(psc_ctx* ctx, int member, const double* u_k, const double* pi_u_k);
int psc_get_gw_k(psc_ctx* ctx, int member, double* u_k, double* pi_u_k);
/* GW spectrum (output.cpp:780-1062) with projection PSC_GW_PROJ_REFERENCE or
 * PSC_GW_PROJ_TT.  out: nbins rows of PSC_GW_SPECTRA_COLS; totals[0..1] =
 * rho_GW / rho, rho_GW (output.cpp:1054-1060). */
int psc_gw_spectrum(psc_ctx* ctx, int member, int projection, double* out, int nbins,
                    double* totals);

/* GW energy spectrum per logarithmic interval, Omega_GW(k) = (1/rho)
 * d rho_GW / d ln k, in nlog bins of ln|n| over [0, ln(sqrt(3) N / 2)]
 * (|n| = k / kIR); rho_GW normalised as in output.cpp:1056-1058.  Has no
 * counterpart in the reference beyond that normalisation.  k_center and omega
 * hold nlog values (nlog <= 1024). */
int psc_gw_omega_log(psc_ctx* ctx, int member, int projection, int nlog, double* k_center,
                     double* omega);

/* Number of FP64 bins of psc_spectra for a lattice of N points per edge. */
int psc_num_bins(int N);

/* ---- adjoint (no counterpart in the reference) -------------------------- */

#define PSC_OBJ_CHI2 0          /* <chi^2> over the lattice at the end of the run */
#define PSC_OBJ_GW_FRACTION 1   /* rho_GW / rho at the end (the totals of psc_gw_spectrum,
                                   output.cpp:1054-1060); needs with_gw              */

typedef struct psc_adjoint_spec {
    int objective;          /* PSC_OBJ_CHI2 or PSC_OBJ_GW_FRACTION                   */
    int nsteps;             /* forward steps of cfg->dt from the loaded state        */
    int checkpoint_every;   /* steps between stored states; each segment is replayed
                               once with recording before it is swept backwards     */
    int member;             /* 0 (single-lattice contexts only)                      */
    int projection;         /* GW objective: PSC_GW_PROJ_REFERENCE or PSC_GW_PROJ_TT */
    int coupling;           /* PSC_COUPLING_*: out[1] is dJ/d(that coupling); the
                               model must provide its derivatives                 */
    double phi0, dphi0, chi0;  /* homogeneous initial values; da/dt(0) = a0 H(phi0)
                                  as initialize.cpp:649-668                         */
    double reserved[4];
} psc_adjoint_spec;

/* Gradient of the objective by the discrete adjoint of the step (reverse
 * kick-drift over the recorded substeps, Friedmann coupling included; with the
 * GW objective also the GW kick/drift and the source adjoint back to the
 * fields), with respect to a coupling of the model (spec->coupling, g2 by
 * default), phi0 (the homogeneous mode and H(phi0),
 * fluctuation modes held fixed) and the initial real-space chi field.  The
 * run is split into segments of checkpoint_every steps: a forward pass stores
 * each segment's start state, then from the last segment backwards the segment
 * is replayed with recording and swept back.  The loaded state is left
 * unchanged.  out[0..5] = J, dJ/d(coupling), dJ/dphi0, forward milliseconds, backward
 * milliseconds (replays and sweeps), device bytes of the adjoint workspace.
 * dchi receives N^3 values dJ/dchi(x) when not NULL.  The workspace
 * (checkpoints, one recorded segment, adjoint arrays) stays allocated for
 * further calls until psc_destroy, psc_set_shape or psc_autotune.  Requires
 * nflds == 2 and batch == 1; the GW objective requires GW evolution and loaded
 * GW data.  No counterpart in the reference. */
int psc_adjoint(psc_ctx* ctx, const psc_adjoint_spec* spec, double* out, double* dchi);

/* ---- launch-shape autotuner (no counterpart in the reference) ----------- */

/* Times every candidate FFT launch shape (radix plan, z columns per block of
 * the strided passes, lines per block of the contiguous passes) for nsteps
 * steps on the loaded state, keeps the fastest, restores the state, and
 * appends "key shape milliseconds" to the cache file named by the environment
 * variable PSC_TUNE_CACHE (default $HOME/.cache/pspectcosmo_amd/tune.txt).
 * report (may be NULL) receives "shape=ms;..." for every candidate.  With
 * pad = 2 the candidates, psc_shape_name, psc_list_shapes and psc_set_shape
 * refer to the N_pad shapes of the padded passes. */
int psc_autotune(psc_ctx* ctx, int nsteps, char* report, int report_len);
/* Name of the launch shape the context uses. */
const char* psc_shape_name(const psc_ctx* ctx);
/* Comma-separated names of the shapes available for the context's N and
 * precision (the default first), and selection of one by name. */
int psc_list_shapes(const psc_ctx* ctx, char* names, int names_len);
int psc_set_shape(psc_ctx* ctx, const char* name);

/* ---- instrumentation --------------------------------------------------- */

/* When on, psc_step brackets every kernel with HIP events on the context's
 * stream and accumulates per-kernel durations (graphs are not used then). */
int psc_set_kernel_timing(psc_ctx* ctx, int on);
/* Per-kernel-kind totals since the last reset: names (comma separated, into
 * names/names_len), total milliseconds, launch counts and the algorithmic bytes
 * each launch moves.  *nkinds receives the number of kinds (<= cap). */
int psc_get_kernel_times(psc_ctx* ctx, char* names, int names_len, double* total_ms,
                         long long* counts, double* bytes_per_launch, int cap, int* nkinds);
int psc_reset_kernel_times(psc_ctx* ctx);
/* Bytes of device memory the context holds. */
int psc_device_bytes(psc_ctx* ctx, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* PSC_H */
