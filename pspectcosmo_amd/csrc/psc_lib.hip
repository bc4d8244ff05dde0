// psc_lib.hip -- host side of the HIP library: device state, kernel
// sequencing of the time step (evolve_VV_k_nopad, evolution.cpp:1131-1198;
// the engines of psc_engine.hpp), hipGraph capture, per-kernel timing and the
// C ABI of include/psc.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "psc_shapes.hpp"
#include "psc_ic_device.hpp"

namespace psc {

// The launch-shape table, assembled from the shape groups in a fixed order.
const std::vector<ShapeEntry>& shape_table() {
    static const std::vector<ShapeEntry> table = [] {
        std::vector<ShapeEntry> v;
        register_shapes_f32_small(v);
        register_shapes_f32_256(v);
        register_shapes_f32_512a(v);
        register_shapes_f32_512b(v);
        register_shapes_f64_small(v);
        register_shapes_f64_256(v);
        register_shapes_f64_512(v);
        register_shapes_r3(v);
        register_shapes_pad768(v);
        return v;
    }();
    return table;
}
#define kShapes (shape_table())
inline int kNumShapes() { return (int)shape_table().size(); }

std::vector<int> shapes_for(int precision, int N) {
    std::vector<int> v;
    for (int i = 0; i < kNumShapes(); ++i)
        if (kShapes[i].precision == precision && kShapes[i].N == N && kShapes[i].make != nullptr) v.push_back(i);
    return v;
}

// Shapes of the padded length NP that run the dealiased stepper of the N lattice.
std::vector<int> pad_shapes_for(int precision, int NP, int N) {
    std::vector<int> v;
    for (int i = 0; i < kNumShapes(); ++i)
        if (kShapes[i].precision == precision && kShapes[i].N == NP && kShapes[i].make_pad != nullptr) {
            std::unique_ptr<PadBase> probe(kShapes[i].make_pad(N));
            if (probe) v.push_back(i);
        }
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

int check_member_only(psc_ctx* c, int member) {
    if (!c) return PSC_ERR_ARG;
    if (member < 0 || member >= c->B) return c->fail(PSC_ERR_ARG, "member index out of range");
    return PSC_OK;
}

// Restores the synchronised momenta and scale factor of every member a call
// left one look-ahead kick ahead (see psc_step): the copies from the sync
// buffers, then the states.  Every function that reads or changes the state
// runs it first.
int ensure_synced(psc_ctx* c) {
    if (!c->ahead) return PSC_OK;
    auto copy = [&](void* dst, const void* src, size_t n, size_t mstride, int arrays, int members) {
        const dim3 grid((unsigned)std::min<size_t>((n + 255) / 256, 4096), (unsigned)(arrays * members));
        if (c->cfg.precision == PSC_FP32)
            hipLaunchKernelGGL(k_sync_copy<float>, grid, dim3(256), 0, c->stream, static_cast<cpx<float>*>(dst),
                               static_cast<const cpx<float>*>(src), c->ms, n, mstride, arrays, members);
        else
            hipLaunchKernelGGL(k_sync_copy<double>, grid, dim3(256), 0, c->stream, static_cast<cpx<double>*>(dst),
                               static_cast<const cpx<double>*>(src), c->ms, n, mstride, arrays, members);
    };
    copy(c->Pp, c->Ps, c->MS, c->MS, 1, c->B);
    if (c->cfg.with_gw) copy(c->PU, c->PUs, c->MS, c->MS, 3, c->B);
    if (c->padeng) copy(c->Gp, c->Gps, (size_t)2 * c->N * c->N, 0, 1, 1);
    hipLaunchKernelGGL(k_sync_state, dim3((c->B + 63) / 64), dim3(64), 0, c->stream, c->ms, c->B);
    HIPCHK(c, hipGetLastError());
    c->ahead = 0;
    return PSC_OK;
}

// A member index check for functions that read or change the state: the
// state is synchronised first.
int check_member(psc_ctx* c, int member) {
    int rc = check_member_only(c, member);
    if (rc) return rc;
    return ensure_synced(c);
}

// The state of newly loaded data (t, a, da/dt; d2a/dt2 follows from it).
MemberState fresh_state(const psc_ctx* c, int member, double t, double a, double ad) {
    MemberState s;
    std::memset(&s, 0, sizeof(s));
    s.t = t;
    s.a = s.a_start = s.a_half = a;
    s.ad = s.b = (c->cfg.expansion == PSC_EXPANSION_STATIC) ? 0.0 : ad;
    s.par = c->models[member];
    s.nflds = c->cfg.nflds;
    s.dt = s.dt_next = c->cfg.dt;
    s.t_out = c->t_out;
    s.dt_cap = c->dt_cap;
    return s;
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
    double rs[kDiagWidth];
    HIPCHK(c, hipMemcpyAsync(rs, c->sums, sizeof(rs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double N3 = (double)c->N3, a = s.a, kIR = c->cfg.kIR;
    for (int t = 0; t < kMaxTerms; ++t) d->Vt[t] = t < Model::num_potential_terms(c->cfg.nflds) ? rs[4 + t] / N3 : 0.0;
    for (int q = 0; q < 2; ++q) {
        d->K[q] = 0.5 * fs[q] / std::pow(a, 6) / (N3 * N3);
        d->G[q] = kIR * kIR * fs[2 + q] / N3 / 2.0 / (a * a) / N3;
        d->mean[q] = rs[q] / N3;
        d->sumsq[q] = rs[2 + q] / N3;
        const double v = d->sumsq[q] - d->mean[q] * d->mean[q];
        d->var[q] = v > 0 ? v : 0.0;
        d->meff[q] = rs[4 + kMaxTerms + q] / N3;
    }
    if (c->cfg.nflds == 1) d->K[1] = d->G[1] = d->mean[1] = d->var[1] = d->sumsq[1] = d->meff[1] = 0.0;
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
    const double K = d.K[0] + d.K[1], G = d.G[0] + d.G[1], V = d.V();
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
        for (int t = 1; t < kMaxTerms; ++t) e.Vt[t] = 0.0;
    }
    *add = kernel_a_value(c, e, a);
    return PSC_OK;
}

// <chi^2> objective: fbar = (P(k) - conj P(-k)) / N^6 on the padded layout.

// Every node of a captured graph must be a kernel node whose grid, block and
// LDS (static + dynamic) fit the device: a memcpy, memset or host node, or a
// launch outside the limits, is refused before the graph is instantiated.
int check_graph(psc_ctx* c, hipGraph_t g, const char* what) {
    size_t n = 0;
    HIPCHK(c, hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    if (n) HIPCHK(c, hipGraphGetNodes(g, nodes.data(), &n));
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, dev));
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType type;
        HIPCHK(c, hipGraphNodeGetType(nodes[i], &type));
        if (type != hipGraphNodeTypeKernel)
            return c->fail(PSC_ERR_DEVICE, std::string(what) + ": captured node " + std::to_string(i) +
                                               " is not a kernel (type " + std::to_string((int)type) + ")");
        hipKernelNodeParams kp;
        HIPCHK(c, hipGraphKernelNodeGetParams(nodes[i], &kp));
        hipFuncAttributes fa;
        HIPCHK(c, hipFuncGetAttributes(&fa, kp.func));
        const size_t threads = (size_t)kp.blockDim.x * kp.blockDim.y * kp.blockDim.z;
        const bool grid_ok = kp.gridDim.x >= 1 && kp.gridDim.y >= 1 && kp.gridDim.z >= 1 &&
                             kp.gridDim.x <= (unsigned)prop.maxGridSize[0] && kp.gridDim.y <= (unsigned)prop.maxGridSize[1] &&
                             kp.gridDim.z <= (unsigned)prop.maxGridSize[2];
        const bool block_ok = threads >= 1 && threads <= (size_t)fa.maxThreadsPerBlock;
        const size_t lds = fa.sharedSizeBytes + kp.sharedMemBytes;
        if (!grid_ok || !block_ok || lds > prop.sharedMemPerBlock || !kp.kernelParams)
            return c->fail(PSC_ERR_DEVICE, std::string(what) + ": captured kernel node " + std::to_string(i) +
                                               " outside the device limits (grid " + std::to_string(kp.gridDim.x) +
                                               "x" + std::to_string(kp.gridDim.y) + "x" +
                                               std::to_string(kp.gridDim.z) + ", block " + std::to_string(threads) +
                                               ", LDS " + std::to_string(lds) + ")");
    }
    return PSC_OK;
}

}  // namespace psc


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
    if (c->cfg.pad) k << "|pad" << c->NP;
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
    if (c->g_last_sync) (void)hipGraphExecDestroy(c->g_last_sync);
    c->g_mid = c->g_last = c->g_last_sync = nullptr;
    c->adj.reset();
}

// Scales of the three snapshot densities: potential (coupling term only with
// two fields), a^6/2 sum pi^2, a^2/2 sum |grad f|^2 (evolution.cpp:7-19,
// 135-137).
void density_scales(const psc_ctx* c, const MemberState& s, double scales[3]) {
    scales[0] = c->cfg.nflds == 2 ? 1.0 : 0.0;
    scales[1] = std::pow(s.a, 6) / 2.0;
    scales[2] = s.a * s.a / 2.0;
}

// The fixed staging array of psc_energy_densities_noalloc: 32 MiB of doubles
// per device, held until the process ends.  Host code never passes `scratch`
// itself to a kernel or a copy (on the host it names a stand-in, not device
// memory): it uses the device address that hipGetSymbolAddress returns, one
// per device, kept in scratch_addr.
constexpr size_t kScratchDoubles = (size_t)4 << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (looked up once, with dev
// current), or nullptr when the lookup fails.
double* scratch_on_device(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_addr[dev]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return nullptr;
        if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        scratch_addr[dev] = static_cast<double*>(p);
    }
    return scratch_addr[dev];
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

const char* psc_model_name(void) { return Model::name; }
const char* psc_model_description(void) { return Model::description; }
int psc_num_potential_terms(int nflds) { return Model::num_potential_terms(nflds); }

int psc_potential_terms(psc_ctx* c, int member, double* out) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!out) return c->fail(PSC_ERR_ARG, "null output");
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    Diag d;
    if ((rc = diagnose(c, member, s, &d))) return rc;
    for (int t = 0; t < Model::num_potential_terms(c->cfg.nflds); ++t) out[t] = d.Vt[t];
    return PSC_OK;
}

int psc_create(const psc_config* cfg, psc_ctx** out) {
    if (!cfg || !out) return PSC_ERR_ARG;
    *out = nullptr;
    const int N = cfg->N;
    if (!psc_ic::lattice_size_ok(N) || cfg->nflds < 1 || cfg->nflds > 2 || cfg->batch < 1 ||
        (cfg->integrator != PSC_VV2 && cfg->integrator != PSC_Y4) ||
        (cfg->precision != PSC_FP64 && cfg->precision != PSC_FP32) || !(cfg->kIR > 0) ||
        !(cfg->fStar > 0) || !(cfg->MPl > 0) || !std::isfinite(cfg->dt) ||
        (cfg->expansion != PSC_EXPANSION_SELF && cfg->expansion != PSC_EXPANSION_STATIC) ||
        (cfg->pad != 0 && cfg->pad != 2) || (cfg->pad == 0 && cfg->n_pad != 0) ||
        (cfg->dt_rule != PSC_DT_FIXED && cfg->dt_rule != PSC_DT_REFERENCE))
        return PSC_ERR_ARG;
    // the dealiased stepper: N_pad (parameters.h:23) larger than N
    // (pspectcosmo.cpp:244-247); 0 selects 2N
    const int NP = cfg->pad ? (cfg->n_pad ? cfg->n_pad : 2 * N) : 0;
    if (cfg->pad && NP <= N) return PSC_ERR_ARG;
    // built: N_pad = 2N and 3N/2 on the padded transform lengths (up to 768),
    // one member, no GW (the reference pads only the scalar force)
    if (cfg->pad && (cfg->batch != 1 || cfg->with_gw || NP > 768 || (NP != 2 * N && 2 * NP != 3 * N)))
        return PSC_ERR_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PSC_ERR_DEVICE;
    std::unique_ptr<psc_ctx> c(new psc_ctx());
    c->cfg = *cfg;
    c->N = N;
    c->NP = NP;
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
        if (cfg->pad) {
            const std::vector<int> pc = pad_shapes_for(cfg->precision, NP, N);
            if (pc.empty()) return PSC_ERR_UNSUPPORTED;
            c->pad_shape = pc[0];
            if (cfg->autotune) {
                const std::string name = tune_lookup(tune_key(c.get()));
                for (int idx : pc)
                    if (name == kShapes[idx].name) c->pad_shape = idx;
            }
            c->padeng = kShapes[c->pad_shape].make_pad(N);
            if (!c->padeng) return PSC_ERR_UNSUPPORTED;
        }
    }
    integrator_weights(*cfg, c->w);
    c->models.assign(c->B, psc_model{1.0, 0.0, 0.0, {0.0, 0.0, 0.0, 0.0, 0.0}});
    c->loaded.assign(c->B, 0);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return PSC_ERR_DEVICE;
    if (hipGetDevice(&c->device) != hipSuccess) return PSC_ERR_DEVICE;
    (void)scratch_on_device(c->device);   // set-up of psc_energy_densities_noalloc's array; its call re-checks
    const size_t arr = (size_t)c->B * c->MS * c->es;
    auto alloc = [&](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        c->dev_bytes += (double)bytes;
        return true;
    };
    // partial-sum sizes: the most blocks any shape of the table can launch
    // (TC >= 8, L >= 4, asserted in Shape), also after a shape switch; the
    // dealiased stepper's x pass runs on its padded grid
    size_t blocks_col = (size_t)(N / 8) * N;
    const size_t blocks_row = (size_t)N * N / 4;
    if (c->padeng) blocks_col = std::max(blocks_col, (size_t)(NP / 8) * NP);
    c->part_len = std::max<size_t>(std::max(blocks_col * kDiagWidth, (size_t)256 * 4),
                                   (size_t)128 * (maxnumbins(N) * 4 + kMaxLogBins));
    c->sums_len = (size_t)maxnumbins(N) * 4 + kMaxLogBins + 64;
    const bool gw = cfg->with_gw != 0;
    if ((gw && (!alloc(&c->U, 3 * arr) || !alloc(&c->PU, 3 * arr) || !alloc(&c->PUs, 3 * arr))) ||
        !alloc(&c->Pf, arr) || !alloc(&c->Pp, arr) || !alloc(&c->Ps, arr) || !alloc(&c->W, gw ? 4 * arr : arr) ||
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
                       !alloc(&c->Gps, (size_t)2 * N * N * c->es) ||
                       !alloc(&c->dcp, c->es) || !alloc(&c->twp, (size_t)NP * c->es)))) {
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
    if (!twiddles(c->tw, N) || (c->padeng && !twiddles(c->twp, NP))) {
        psc_destroy(c.release());
        return PSC_ERR_DEVICE;
    }
    std::vector<MemberState> init(c->B);
    for (auto& s : init) {
        std::memset(&s, 0, sizeof(s));
        s.a = s.a_start = s.a_half = 1.0;
        s.par = c->models[0];
        s.nflds = cfg->nflds;
        s.dt = s.dt_next = cfg->dt;
        s.t_out = s.dt_cap = 1e300;
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
    void* ptrs[] = {c->Pf, c->Pp, c->Ps, c->W, c->tw, c->U, c->PU, c->PUs, c->ms, c->partV, c->partKG, c->partV2,
                    c->partKG2, c->part, c->sums, c->Wp, c->Rp, c->Gf, c->Gp, c->Gps, c->dcp, c->twp};
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
    s.par = *m;
    return write_state(c, member, &s);
}

// The dealiased stepper's half spectra in its own form: in the kz = 0 and
// kz = N/2 planes each field's value splits into the Hermitian part
// (h(n) + conj h(-n)) / 2, which the packed spectrum holds, and the
// anti-Hermitian part (h(n) - conj h(-n)) / 2, which goes to the plane ghost
// Gamma = A_phi + i A_chi (the parts psc_get_k adds back).  `herm` receives
// the half spectra with their Hermitian parts, G the ghost (2 N^2 complex,
// plane-major, then x, then y).
void split_plane_ghosts(int N, int nflds, const double* hk, std::vector<double>& herm, std::vector<double>& G) {
    const int Nh = N / 2 + 1;
    const size_t nk = (size_t)N * N * Nh;
    herm.assign(hk, hk + 2 * nk * nflds);
    G.assign((size_t)2 * 2 * N * N, 0.0);
    for (int pl = 0; pl < 2; ++pl) {
        const int k = pl ? N / 2 : 0;
        for (int x = 0; x < N; ++x)
            for (int y = 0; y < N; ++y) {
                const size_t at = ((size_t)x * N + y) * Nh + k;
                const size_t mat = ((size_t)((N - x) % N) * N + (N - y) % N) * Nh + k;
                double A[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
                for (int q = 0; q < nflds; ++q) {
                    const double* h = hk + 2 * nk * q;
                    const double hr = h[2 * at], hi = h[2 * at + 1], mr = h[2 * mat], mi = -h[2 * mat + 1];
                    herm[2 * nk * q + 2 * at] = 0.5 * (hr + mr);
                    herm[2 * nk * q + 2 * at + 1] = 0.5 * (hi + mi);
                    A[q][0] = 0.5 * (hr - mr);
                    A[q][1] = 0.5 * (hi - mi);
                }
                const size_t g = (size_t)pl * N * N + (size_t)x * N + y;
                G[2 * g] = A[0][0] - A[1][1];       // A_phi + i A_chi
                G[2 * g + 1] = A[0][1] + A[1][0];
            }
    }
}

// Upload a plane ghost (2 N^2 complex FP64) in the field precision.
int upload_ghost(psc_ctx* c, void* dst, const std::vector<double>& G) {
    if (c->cfg.precision == PSC_FP32) {
        std::vector<float> tmp(G.size());
        for (size_t i = 0; i < G.size(); ++i) tmp[i] = (float)G[i];
        HIPCHK(c, hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(dst, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return PSC_OK;
}

int psc_load_ic_k(psc_ctx* c, int member, const double* f_k, const double* pi_k, double a, double ad) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!f_k || !pi_k || !(a > 0)) return c->fail(PSC_ERR_ARG, "bad initial data");
    std::vector<double> full;
    // FP32: the homogeneous mode's FP64 remainders beside its FP32 values
    double lo[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    auto remainder = [&](double (&r)[2]) {
        if (c->cfg.precision != PSC_FP32) return;
        for (int i = 0; i < 2; ++i) r[i] = full[i] - (double)(float)full[i];
    };
    // the dealiased stepper keeps the anti-Hermitian parts of the kz = 0, N/2
    // planes in its ghosts (a state read by psc_get_k loads back unchanged)
    std::vector<double> herm, gf, gp;
    if (c->padeng) split_plane_ghosts(c->N, c->cfg.nflds, f_k, herm, gf);
    pack_full(c->N, c->cfg.nflds, c->padeng ? herm.data() : f_k, full);
    remainder(lo[0]);
    if ((rc = upload(c, c->Pf, member, full))) return rc;
    if (c->padeng) split_plane_ghosts(c->N, c->cfg.nflds, pi_k, herm, gp);
    pack_full(c->N, c->cfg.nflds, c->padeng ? herm.data() : pi_k, full);
    remainder(lo[1]);
    if ((rc = upload(c, c->Pp, member, full))) return rc;
    if ((rc = reset_ghosts(c))) return rc;
    if (c->padeng && ((rc = upload_ghost(c, c->Gf, gf)) || (rc = upload_ghost(c, c->Gp, gp)))) return rc;
    MemberState s = fresh_state(c, member, 0.0, a, ad);
    for (int i = 0; i < 2; ++i) {
        s.dc_f[i] = lo[0][i];
        s.dc_p[i] = lo[1][i];
    }
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
    const psc_ic::Pot V(cfg.nflds, mdl);
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
    // kept modes per field (all of them without a kcutoff) and, with a cut,
    // the interior row offsets, copied into W for the kernel to read
    const psc_ic::Cut cut = psc_ic::make_cut(cfg, ic->kcutoff);
    psc_ic::KeptIndex kf;
    std::vector<uint64_t> rowpre_f, rowpre_gw;
    const unsigned long long M = psc_ic::build_kept_index(N, true, cut, kf, rowpre_f);
    const std::vector<uint64_t>* rowpre = &rowpre_f;
    if (!gw) {
        k->norm = cfg.omegaStar / cfg.fStar * std::pow(N / dx, 1.5) / std::sqrt(2.0);
        const double f0[2] = {ic->f0[0], cfg.nflds > 1 ? ic->f0[1] : 0.0};
        V.meff(k->meff, f0);
        for (int q = 0; q < cfg.nflds; ++q) {
            k->first[q] = 4 * (unsigned long long)q * M;
            k->f0[q] = ic->f0[q];
            k->fd0[q] = ic->fd0[q];
        }
        k->kept = kf;
    } else {
        k->norm = cfg.MPl * cfg.omegaStar / (cfg.fStar * cfg.fStar) * std::pow(N / dx, 1.5);
        k->first[0] = 4 * M * (unsigned long long)cfg.nflds;   // after every kept scalar mode's draws
        psc_ic::build_kept_index(N, false, cut, k->kept, rowpre_gw);
        rowpre = &rowpre_gw;
    }
    k->kept.rowpre = nullptr;
    if (cut.on()) {   // W is free between steps; the table is N (N/2 - 1) offsets
        HIPCHK(c, hipMemcpyAsync(c->W, rowpre->data(), rowpre->size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        k->kept.rowpre = static_cast<const uint64_t*>(c->W);
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
    MemberState s = fresh_state(c, member, 0.0, sc[0], sc[1]);
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
    MemberState s = fresh_state(c, member, t, a, ad);
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
    int rc = check_member_only(c, member);   // no synchronisation: a read of t and a
    if (rc) return rc;
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    if (s.ahead) s.a = s.a_start;   // the synchronised scale factor of a member ahead
    out[0] = s.t;
    out[1] = s.a;
    out[2] = s.ad;
    out[3] = s.add;
    if (s.a < 0) return c->fail(PSC_ERR_NEGATIVE_A, "scale factor is negative");
    return PSC_OK;
}

int psc_get_step_count(psc_ctx* c, int member, long long* steps) {
    int rc = check_member_only(c, member);
    if (rc) return rc;
    if (!steps) return c->fail(PSC_ERR_ARG, "null output");
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    *steps = (long long)s.steps;
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
    // FP32: the homogeneous mode is its FP32 value plus the FP64 remainder
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    if (f_k) {
        if ((rc = download(c, c->Pf, member, full))) return rc;
        full[0] += s.dc_f[0];
        full[1] += s.dc_f[1];
        unpack_half(c->N, c->cfg.nflds, full.data(), f_k);
        if (c->padeng && (rc = add_ghost(c->Gf, f_k))) return rc;
    }
    if (pi_k) {
        if ((rc = download(c, c->Pp, member, full))) return rc;
        full[0] += s.dc_p[0];
        full[1] += s.dc_p[1];
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
    double scales[3];
    density_scales(c, s, scales);
    for (int kind = 0; kind < 3; ++kind) {
        double* dst = outs[kind];
        if (!dst && !tot) continue;
        if (!dst) {
            tmp.resize(n3);
            dst = tmp.data();
        }
        if ((rc = c->eng->density(c, member, kind, D, scales[kind], 0, c->N))) return rc;
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

int psc_energy_densities_noalloc(psc_ctx* c, int member, double* pot, double* kin, double* grad, double* tot) {
    int rc = check_member(c, member);
    if (rc) return rc;
    if (tot && (!pot || !kin || !grad))
        return c->fail(PSC_ERR_ARG, "tot needs pot, kin and grad: this call has no host staging");
    const size_t row = (size_t)c->N * c->N;   // one y row of every x plane, in doubles
    if (row > kScratchDoubles)
        return c->fail(PSC_ERR_UNSUPPORTED, "N^2 doubles exceed the 32 MiB staging array");
    double* S = scratch_on_device(c->device);
    if (!S) return c->fail(PSC_ERR_DEVICE, "hipGetSymbolAddress of the staging array failed");
    MemberState s;
    if ((rc = read_state(c, member, &s))) return rc;
    double scales[3];
    density_scales(c, s, scales);
    const int N = c->N;
    const int rows = (int)std::min<size_t>((size_t)N, kScratchDoubles / row);
    double* outs[3] = {pot, kin, grad};
    for (int kind = 0; kind < 3; ++kind) {
        double* dst = outs[kind];
        if (!dst) continue;
        // y rows o0 .. o0 + r - 1 of every x plane: N blocks of r N doubles,
        // one per plane, placed at (x N + o0) N in the output
        for (int o0 = 0; o0 < N; o0 += rows) {
            const int r = std::min(rows, N - o0);
            if ((rc = c->eng->density(c, member, kind, S, scales[kind], o0, r))) return rc;
            const size_t w = (size_t)r * N * sizeof(double);
            HIPCHK(c, hipMemcpy2DAsync(dst + (size_t)o0 * N, row * sizeof(double), S, w, w, (size_t)N,
                                       hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (tot) {  // tot = pot + kin + grad in that order (output.cpp:1623)
            if (kind == 0)
                std::memcpy(tot, dst, c->N3 * sizeof(double));
            else
                for (size_t i = 0; i < c->N3; ++i) tot[i] += dst[i];
        }
    }
    return PSC_OK;
}

int psc_set_dt(psc_ctx* c, double dt) {
    if (!c) return PSC_ERR_ARG;
    if (!(dt > 0.0) || !std::isfinite(dt)) return c->fail(PSC_ERR_ARG, "dt must be positive");
    int rc = ensure_synced(c);   // a look-ahead kick applied with the old dt is undone
    if (rc) return rc;
    c->cfg.dt = dt;
    for (int m = 0; m < c->B; ++m) {
        MemberState s;
        if ((rc = read_state(c, m, &s))) return rc;
        s.dt = s.dt_next = dt;
        if ((rc = write_state(c, m, &s))) return rc;
    }
    return PSC_OK;
}

int psc_set_output_times(psc_ctx* c, double t_out, double dt_cap) {
    if (!c) return PSC_ERR_ARG;
    if (!(dt_cap > 0.0) || std::isnan(t_out)) return c->fail(PSC_ERR_ARG, "bad output times");
    int rc = ensure_synced(c);   // a look-ahead kick applied with a dt of the old targets is undone
    if (rc) return rc;
    c->t_out = t_out;
    c->dt_cap = dt_cap;
    for (int m = 0; m < c->B; ++m) {
        MemberState s;
        if ((rc = read_state(c, m, &s))) return rc;
        s.t_out = t_out;
        s.dt_cap = dt_cap;
        if ((rc = write_state(c, m, &s))) return rc;
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
    for (int t = 0; t < kMaxTerms; ++t) {
        if (t < 2) out[4 + t] = d.Vt[t];
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
    // <m_eff^2> of the field: effective_mass (model.h:96-98) averaged over
    // the lattice as output.cpp:569-580
    Diag dg;
    if ((rc = diagnose(c, member, s, &dg))) return rc;
    const double msq = dg.meff[fld];
    const double N3 = (double)c->N3;
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
    int rc0 = ensure_synced(c);   // the snapshot is of the synchronised state
    if (rc0) return rc0;
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
        if (!save) c->ahead = 0;
        return PSC_OK;
    };
    int rc = copy(true);
    if (rc) { cleanup(); return rc; }
    // the dealiased stepper's time is in its padded passes: tune those
    const std::vector<int> cand =
        pad ? pad_shapes_for(c->cfg.precision, c->NP, c->N) : shapes_for(c->cfg.precision, c->N);
    const int saved_timing = c->timing;
    c->timing = 0;
    double best_ms = 1e300;
    int best = pad ? c->pad_shape : c->shape;
    auto adopt = [&](int idx) {
        drop_graphs(c);
        if (pad) {
            delete c->padeng;
            c->padeng = kShapes[idx].make_pad(c->N);
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
    for (int idx : c->padeng ? pad_shapes_for(c->cfg.precision, c->NP, c->N) : shapes_for(c->cfg.precision, c->N)) {
        if (!s.empty()) s += ",";
        s += kShapes[idx].name;
    }
    std::strncpy(names, s.c_str(), names_len - 1);
    names[names_len - 1] = 0;
    return PSC_OK;
}

int psc_set_shape(psc_ctx* c, const char* name) {
    if (!c || !name) return PSC_ERR_ARG;
    if (c->padeng) {   // the dealiased stepper's padded shape
        for (int idx : pad_shapes_for(c->cfg.precision, c->NP, c->N))
            if (std::strcmp(name, kShapes[idx].name) == 0) {
                HIPCHK(c, hipStreamSynchronize(c->stream));
                drop_graphs(c);
                delete c->padeng;
                c->pad_shape = idx;
                c->padeng = kShapes[idx].make_pad(c->N);
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
    if (!has_coupling<Model>(sp->coupling))
        return c->fail(PSC_ERR_UNSUPPORTED, std::string("model ") + Model::name +
                                                " has no derivatives for coupling " + std::to_string(sp->coupling));
    if (!c->loaded[0]) return c->fail(PSC_ERR_STATE, "no initial data");
    if (c->cfg.dt_rule != PSC_DT_FIXED) return c->fail(PSC_ERR_UNSUPPORTED, "the adjoint takes fixed steps");
    int rs = ensure_synced(c);
    if (rs) return rs;
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
