// psc_run.cpp -- command-line driver over the C ABI: the main loop of the
// reference program (main, pspectcosmo.cpp:237-705) for the spectral stepper
// (evolve_VV_k_nopad), writing the reference's text outputs (output.cpp):
//   average_scalar_<f>.txt, average_scale_factor.txt, average_energies.txt,
//   average_energy_conservation.txt          every `freq` (save_Freq, output.cpp:1831-1840)
//   spectra_scalar_<f>.txt, spectra_scalar_<f>_evolution.txt,
//   average_spectra_times.txt, spectra_gws.txt, energy_gws.txt
//                                            every `infreq` (save_Infreq, output.cpp:1842-1848)
//   info.txt                                 at the end (output_parameters, output.cpp:1750)
// snapshots every `snap` (save_Snap, output.cpp:1850-1853): HDF5 files
// Snapshots/data_<n>_snapshot.h5 with /ScalarField_<f> and, with
// snap_energy = 1 (SNAPSHOTS_ENERGY, parameters.h:14), /PotentialEnergy,
// /KineticEnergy, /GradientEnergy, /TotalEnergy, all N x N x N native doubles
// (snapshots, output.cpp:1541-1648); the HDF5 C library is loaded when the
// first snapshot is due (snap_format = raw writes plain float64 files instead),
// and the lattice image grid.img (checkpoint, output.cpp:1652-1703) with
// every spectra output and at the end when checkpoint = 1; continue_run = 1
// resumes from it (initialize.cpp:590-616).
//
// Parameters (names follow parameters.h) come from an optional file of
// `key = value` lines and from `key=value` arguments; `psc_run --help` lists
// them.  Two time-step modes:
//   dt = <value>  fixed step; output every freq_steps / infreq_steps steps;
//   dt = auto     the reference's rule dt = 2 / (kIR sqrt(3) N / 2 / a) / 10,
//                 capped at tOutputFreq and clamped to the next output time
//                 (pspectcosmo.cpp:619-627), outputs at t within 1e-10 of the
//                 output times (pspectcosmo.cpp:652-697).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/psc.h"
#include "psc_hdf5.hpp"

namespace {

const int kFreqPrec = 3, kInfreqPrec = 2, kSetw = 15, kKirPrec = 1;   // parameters.h:29,51-55

struct Params {
    std::map<std::string, std::string> kv;
    std::string get(const std::string& k, const std::string& def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : it->second;
    }
    double num(const std::string& k, double def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : std::strtod(it->second.c_str(), nullptr);
    }
    std::vector<double> list(const std::string& k, std::vector<double> def) const {
        auto it = kv.find(k);
        if (it == kv.end()) return def;
        std::vector<double> v;
        std::string s = it->second;
        for (char& ch : s)
            if (ch == ',') ch = ' ';
        std::istringstream in(s);
        double x;
        while (in >> x) v.push_back(x);
        return v;
    }
};

const char* kKeys[][2] = {
    {"N", "points per edge: 2^k or 3 * 2^k, 16..512"},
    {"nflds", "1 or 2"},
    {"kIR", "infrared momentum; L = 2 pi / kIR"},
    {"dt", "time step, or `auto` for the reference's adaptive rule"},
    {"tf", "final time (default: nsteps * dt)"},
    {"nsteps", "fixed-dt runs: number of steps"},
    {"tOutputFreq", "auto dt: interval of the frequent outputs"},
    {"tOutputInfreq", "auto dt: interval of the spectra"},
    {"tSnapFreq", "auto dt: interval of the snapshots (0: none)"},
    {"freq_steps", "fixed dt: steps between frequent outputs"},
    {"infreq_steps", "fixed dt: steps between spectra"},
    {"snap_steps", "fixed dt: steps between snapshots (0: none)"},
    {"integrator", "VV (2) or Y4 (4)"},
    {"pad", "0: evolve_VV_k_nopad; 2: the dealiased stepper (PAD_FLAG)"},
    {"N_pad", "padded lattice of the dealiased stepper: 2N or 3N/2 (0: 2N)"},
    {"yoshida_exact", "1: Yoshida weights from 2^(1/3); 0: the reference's digits"},
    {"precision", "fp64 or fp32"},
    {"expansion", "0 no expansion, 2 self-consistent"},
    {"alpha", "time variable exponent (0: cosmic time)"},
    {"fStar", "field unit"},
    {"omegaStar", "time unit"},
    {"MPl", "Planck mass"},
    {"m2", "mass coupling of the model (psc_model.m2)"},
    {"g2", "phi^2 chi^2 coupling (psc_model.g2)"},
    {"lambda", "phi^4 coupling (psc_model.lambda)"},
    {"couplings", "further couplings c[0..4] of a user model, comma separated"},
    {"initfield", "homogeneous field values, comma separated"},
    {"initderivs", "homogeneous field derivatives, comma separated"},
    {"seed", "Park-Miller seed"},
    {"kcutoff", "low-pass of the initial modes: k above it start at zero (0: none)"},
    {"ic", "host (the host generator) or device (the same modes generated on the GPU)"},
    {"with_gw", "1: evolve the 6 GW components"},
    {"gw_projection", "ref (the reference's one-sided projection) or tt"},
    {"use_graph", "1: replay captured steps"},
    {"autotune", "1: time the launch shapes before the run"},
    {"dir", "output directory (default output)"},
    {"snap_format", "h5 (the reference's HDF5 snapshots) or raw (float64 files)"},
    {"snap_energy", "1: snapshots also hold the energy densities (SNAPSHOTS_ENERGY)"},
    {"hdf5_lib", "path of libhdf5.so (default: the loader's search, then /opt/conda/lib)"},
    {"checkpoint", "1: write <dir>/grid.img with the spectra and at the end"},
    {"continue_run", "1: resume from <dir>/grid.img when it exists"},
    {"store_lattice_times", "times at which the lattice image moves to a new file grid<T>.img, comma separated"},
};

void usage() {
    std::printf("usage: psc_run [param-file] [key=value ...] [--dry-run]\n\nkeys:\n");
    for (auto& k : kKeys) std::printf("  %-15s %s\n", k[0], k[1]);
}

bool parse_line(const std::string& line, Params& p) {
    std::string s = line.substr(0, line.find('#'));
    const size_t eq = s.find('=');
    if (eq == std::string::npos) return false;
    auto trim = [](std::string x) {
        const size_t a = x.find_first_not_of(" \t\r\n"), b = x.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
    };
    const std::string k = trim(s.substr(0, eq)), v = trim(s.substr(eq + 1));
    if (k.empty()) return false;
    p.kv[k] = v;
    return true;
}

bool known_key(const std::string& k) {
    for (auto& e : kKeys)
        if (k == e[0]) return true;
    return false;
}

struct Run {
    psc_config cfg{};
    psc_model model{};
    psc_ic_params ic{};
    psc_ctx* ctx = nullptr;
    std::string dir = "output";
    bool auto_dt = false;
    int gw_proj = PSC_GW_PROJ_REFERENCE;
    int nbins = 0;
    std::ofstream fields[2], sf, energy, conservation, spectra[2], spectra_ev[2], spectra_times, gw_spec, gw_energy;
    double total_initial = 0.0;
    bool first_energy = true;

    int check(int rc, const char* what) {
        if (rc != PSC_OK) {
            std::fprintf(stderr, "psc_run: %s failed (%s): %s\n", what, psc_status_string(rc),
                         ctx ? psc_last_error(ctx) : "");
            std::exit(rc == PSC_ERR_NONFINITE || rc == PSC_ERR_NEGATIVE_A ? 1 : 2);
        }
        return rc;
    }

    std::string path(const std::string& name) const { return dir + "/" + name + ".txt"; }

    void open_outputs() {
        for (int f = 0; f < cfg.nflds; ++f) {
            fields[f].open(path("average_scalar_" + std::to_string(f)), std::ios::out | std::ios::app);
            spectra[f].open(path("spectra_scalar_" + std::to_string(f)), std::ios::out | std::ios::app);
            spectra_ev[f].open(path("spectra_scalar_" + std::to_string(f) + "_evolution"),
                               std::ios::out | std::ios::app);
            // header row of bin momenta (output.cpp:527-533)
            for (int i = 0; i < nbins + 1; ++i)
                spectra_ev[f] << std::fixed << std::setprecision(kKirPrec) << cfg.kIR * i << " ";
            spectra_ev[f] << std::endl;
        }
        if (cfg.expansion == PSC_EXPANSION_SELF)
            sf.open(path("average_scale_factor"), std::ios::out | std::ios::app);
        energy.open(path("average_energies"), std::ios::out | std::ios::app);
        conservation.open(path("average_energy_conservation"), std::ios::out | std::ios::app);
        spectra_times.open(path("average_spectra_times"), std::ios::out | std::ios::app);
        if (cfg.with_gw) {
            gw_spec.open(path("spectra_gws"), std::ios::out | std::ios::app);
            gw_energy.open(path("energy_gws"), std::ios::out | std::ios::app);
        }
    }

    // save_Freq: meansvars, scale, energy (output.cpp:380-486,1069-1187)
    void save_freq() {
        double sc[4];
        check(psc_get_scalars(ctx, 0, sc), "psc_get_scalars");
        const double t = sc[0], a = sc[1], ad = sc[2], add = sc[3];
        double m[12];
        check(psc_means(ctx, 0, m), "psc_means");
        for (int f = 0; f < cfg.nflds; ++f) {
            const double* o = m + 6 * f;
            fields[f] << std::fixed << std::setprecision(kFreqPrec) << t << std::scientific << std::setprecision(6)
                      << std::setw(kSetw) << o[0] << std::setw(kSetw) << o[1] << std::setw(kSetw) << o[2]
                      << std::setw(15) << o[3] << std::setw(kSetw) << o[4] << std::setw(kSetw) << o[5]
                      << std::endl;
        }
        if (cfg.expansion == PSC_EXPANSION_SELF) {
            sf << std::fixed << std::setprecision(6) << t << " " << a << " " << ad << " " << ad / a << " " << add
               << " " << std::scientific << std::setprecision(6) << std::setw(kSetw)
               << 1 / (std::pow(a, -cfg.alpha) * ad) << " " << std::fixed << std::setprecision(6)
               << 1 + cfg.alpha - a * add / (ad * ad) << std::endl;
        }
        double e[PSC_ENERGY_LEN];
        check(psc_energies(ctx, 0, e), "psc_energies");
        const int nterms = psc_num_potential_terms(cfg.nflds);
        double vt[8] = {0};
        check(psc_potential_terms(ctx, 0, vt), "psc_potential_terms");
        std::vector<double> vals;
        for (int f = 0; f < cfg.nflds; ++f) {
            vals.push_back(e[2 * f]);
            vals.push_back(e[2 * f + 1]);
        }
        for (int k = 0; k < nterms; ++k) vals.push_back(vt[k]);
        const double total = e[6];
        vals.push_back(total);
        energy << std::fixed << std::setprecision(kFreqPrec) << t;
        for (double v : vals) energy << std::scientific << std::setprecision(6) << " " << v;
        energy << std::endl;
        if (first_energy) {
            total_initial = total;
            first_energy = false;
        }
        if (cfg.expansion == PSC_EXPANSION_STATIC) {
            conservation << std::fixed << std::setprecision(kFreqPrec) << t << " " << std::scientific
                         << std::setprecision(6) << total / total_initial << std::endl;
        } else {
            const double f2m = (cfg.fStar / cfg.MPl) * (cfg.fStar / cfg.MPl);
            const double LHS = ad * ad, RHS = f2m * std::pow(a, 2 * (cfg.alpha + 1)) * total / 3.;
            conservation << std::fixed << std::setprecision(kFreqPrec) << t << " " << std::scientific
                         << std::setprecision(8) << std::fabs(LHS - RHS) / std::fabs(LHS + RHS) << " " << std::fixed
                         << std::setprecision(6) << LHS << " " << RHS << std::endl;
        }
    }

    // save_Infreq: spectra (output.cpp:490-705, GW 780-1062)
    void save_infreq() {
        double sc[4];
        check(psc_get_scalars(ctx, 0, sc), "psc_get_scalars");
        const double t = sc[0];
        const double N = cfg.N, N3 = N * N * N, L = 2 * M_PI / cfg.kIR, dx = L / N;
        std::vector<double> rows((size_t)nbins * PSC_SPECTRA_COLS);
        for (int f = 0; f < cfg.nflds; ++f) {
            check(psc_spectra(ctx, 0, f, rows.data(), nbins), "psc_spectra");
            for (int b = 0; b < nbins; ++b) {
                const double* r = &rows[(size_t)b * PSC_SPECTRA_COLS];
                spectra[f] << std::fixed << std::setprecision(kKirPrec) << r[0] << " " << std::scientific
                           << std::setprecision(6) << std::setw(kSetw) << r[1] << " " << std::setw(kSetw) << r[2]
                           << " " << std::setw(kSetw) << r[3] << " " << (long long)r[4] << std::endl;
            }
            spectra[f] << std::endl;
            // amplitude per bin (N/dx) omegaStar L^{3/2} / N^3 sqrt(<|f_k|^2>) (output.cpp:693-700)
            spectra_ev[f] << std::fixed << std::setprecision(6) << t << " ";
            for (int b = 0; b < nbins; ++b) {
                const double* r = &rows[(size_t)b * PSC_SPECTRA_COLS];
                const int i = b + 1;
                const double f2 = r[4] > 0 ? r[1] * N3 * N3 / (i * r[4]) : 0.0;
                spectra_ev[f] << std::scientific << std::setprecision(6) << std::setw(kSetw)
                              << (N / dx) * cfg.omegaStar * std::pow(L, 1.5) / N3 * std::sqrt(f2) << " ";
            }
            spectra_ev[f] << std::endl;
        }
        spectra_times << std::fixed << std::setprecision(kInfreqPrec) << t << std::endl;
        if (cfg.with_gw) {
            std::vector<double> g((size_t)nbins * PSC_GW_SPECTRA_COLS);
            double tot[2];
            check(psc_gw_spectrum(ctx, 0, gw_proj, g.data(), nbins, tot), "psc_gw_spectrum");
            for (int b = 0; b < nbins; ++b) {
                const double* r = &g[(size_t)b * PSC_GW_SPECTRA_COLS];
                gw_spec << std::fixed << std::setprecision(kKirPrec) << r[0] << " " << std::scientific
                        << std::setprecision(6) << std::setw(kSetw) << r[1] << " " << std::scientific
                        << std::setprecision(6) << std::setw(kSetw) << r[2] << " " << std::scientific
                        << std::setprecision(6) << std::setw(kSetw) << r[3] << " " << (long long)r[4] << std::endl;
            }
            gw_spec << std::endl;
            gw_energy << std::fixed << std::setprecision(kInfreqPrec) << t << " " << std::scientific
                      << std::setprecision(6) << tot[0] << " " << tot[1] << std::endl;
        }
    }

    // Snapshot (snapshots, output.cpp:1541-1648): the real-space fields and,
    // with snap_energy, the four energy densities, as N^3 doubles x-major.
    // HDF5 by default; snap_format = raw writes each dataset as a little-endian
    // float64 file plus a text index.
    bool snap_h5 = true, snap_energy = true;
    std::string hdf5_lib;
    psc_io::Hdf5 h5;
    void save_snap(int number) {
        const std::string sdir = dir + "/Snapshots";
        mkdir(sdir.c_str(), 0755);
        const size_t n3 = (size_t)cfg.N * cfg.N * cfg.N;
        std::vector<double> f(n3 * cfg.nflds);
        check(psc_get_fields(ctx, 0, f.data(), nullptr), "psc_get_fields");
        std::vector<std::pair<std::string, const double*>> sets;
        for (int q = 0; q < cfg.nflds; ++q) sets.push_back({"/ScalarField_" + std::to_string(q), f.data() + q * n3});
        std::vector<double> en;
        if (snap_energy) {
            en.resize(4 * n3);
            check(psc_energy_densities(ctx, 0, en.data(), en.data() + n3, en.data() + 2 * n3, en.data() + 3 * n3),
                  "psc_energy_densities");
            const char* names[4] = {"/PotentialEnergy", "/KineticEnergy", "/GradientEnergy", "/TotalEnergy"};
            for (int i = 0; i < 4; ++i) sets.push_back({names[i], en.data() + i * n3});
        }
        const std::string base = sdir + "/data_" + std::to_string(number) + "_snapshot";
        if (snap_h5) {
            std::string err = h5.load(hdf5_lib);
            if (err.empty()) err = h5.write(base + ".h5", cfg.N, sets);
            if (!err.empty()) {
                std::fprintf(stderr, "psc_run: snapshot: %s\n", err.c_str());
                std::exit(2);
            }
            return;
        }
        double sc[4];
        check(psc_get_scalars(ctx, 0, sc), "psc_get_scalars");
        std::ofstream idx(base + ".txt");
        idx << "t " << std::setprecision(17) << sc[0] << "\na " << sc[1] << "\nN " << cfg.N << "\n";
        for (auto& d : sets) {
            const std::string name = d.first.substr(1);
            std::ofstream out(base + "." + name + ".f64", std::ios::binary);
            out.write(reinterpret_cast<const char*>(d.second), n3 * sizeof(double));
            idx << d.first << " float64 " << cfg.N << "x" << cfg.N << "x" << cfg.N << " " << name << ".f64\n";
        }
    }

    // Lattice image in the reference's layout (checkpoint, output.cpp:1690-1697):
    // int run_number, double t, a, da/dt, then f and pi, nflds x N^3 doubles each.
    // Image file switching at store_lattice_times (checkpoint,
    // output.cpp:1653-1683): with no times (or a first time of 0) every image
    // overwrites grid.img; otherwise the images go to grid<T>.img, T the
    // integer part of the listed time they run up to, and the first image
    // written after t passes that time moves on to the next name (one move
    // per image), the last file being named for tf.
    std::vector<double> lattice_times;
    size_t lattice_current = 0;
    bool lattice_first = true;
    std::string lattice_name;
    double tf = 0.0;

    std::string grid_name(double t) {
        const bool multi = !lattice_times.empty() && lattice_times[0] != 0.0;
        if (!multi) return dir + "/grid.img";
        if (lattice_first) {
            lattice_first = false;
            lattice_name = dir + "/grid" + std::to_string((int)lattice_times[0]) + ".img";
        } else if (lattice_current < lattice_times.size() && t > lattice_times[lattice_current]) {
            ++lattice_current;
            const double upto = lattice_current < lattice_times.size() ? lattice_times[lattice_current] : tf;
            lattice_name = dir + "/grid" + std::to_string((int)upto) + ".img";
        }
        return lattice_name;
    }

    // Beside each image, <image>.k holds the k-space state the reference's
    // image cannot: the FP64 half spectra of f and pi (with the dealiased
    // stepper's plane parts), d2a/dt2, and with GW the u_ij and pi_u_ij.  It
    // names the header of its image (run number, t, a, da/dt) and is read on
    // continue_run only when the image has that header, so an image replaced
    // by hand is read as the reference reads it.
    static constexpr char kSideMagic[8] = {'P', 'S', 'C', 'K', 'I', 'M', 'G', '1'};
    size_t half_len() const { return (size_t)cfg.N * cfg.N * (cfg.N / 2 + 1) * 2; }

    void write_grid(int run_number) {
        const size_t n3 = (size_t)cfg.N * cfg.N * cfg.N;
        std::vector<double> f(n3 * cfg.nflds), p(n3 * cfg.nflds);
        check(psc_get_fields(ctx, 0, f.data(), p.data()), "psc_get_fields");
        double sc[8];
        check(psc_get_scalars(ctx, 0, sc), "psc_get_scalars");
        const std::string name = grid_name(sc[0]);
        std::ofstream g(name + ".tmp", std::ios::binary);
        g.write(reinterpret_cast<const char*>(&run_number), sizeof(run_number));
        g.write(reinterpret_cast<const char*>(&sc[0]), sizeof(double));
        g.write(reinterpret_cast<const char*>(&sc[1]), sizeof(double));
        g.write(reinterpret_cast<const char*>(&sc[2]), sizeof(double));
        g.write(reinterpret_cast<const char*>(f.data()), f.size() * sizeof(double));
        g.write(reinterpret_cast<const char*>(p.data()), p.size() * sizeof(double));
        g.close();
        // the k-space state
        std::vector<double> fk(half_len() * 2), pk(half_len() * 2);
        check(psc_get_k(ctx, 0, fk.data(), pk.data()), "psc_get_k");
        std::ofstream k(name + ".k.tmp", std::ios::binary);
        const int hdr[6] = {cfg.N, cfg.nflds, cfg.pad, cfg.pad ? (cfg.n_pad ? cfg.n_pad : 2 * cfg.N) : 0, cfg.with_gw,
                            run_number};
        k.write(kSideMagic, sizeof(kSideMagic));
        k.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
        k.write(reinterpret_cast<const char*>(sc), 4 * sizeof(double));   // t, a, da/dt, d2a/dt2
        k.write(reinterpret_cast<const char*>(fk.data()), half_len() * cfg.nflds * sizeof(double));
        k.write(reinterpret_cast<const char*>(pk.data()), half_len() * cfg.nflds * sizeof(double));
        if (cfg.with_gw) {
            std::vector<double> u(6 * half_len()), pu(6 * half_len());
            check(psc_get_gw_k(ctx, 0, u.data(), pu.data()), "psc_get_gw_k");
            k.write(reinterpret_cast<const char*>(u.data()), u.size() * sizeof(double));
            k.write(reinterpret_cast<const char*>(pu.data()), pu.size() * sizeof(double));
        }
        k.close();
        std::rename((name + ".k.tmp").c_str(), (name + ".k").c_str());
        std::rename((name + ".tmp").c_str(), name.c_str());
    }

    // The k-space state of grid.img, when <dir>/grid.img.k belongs to it;
    // sets gw_restored when it also held the GW state.
    bool read_side(int run_number, double t0, double a, double ad) {
        std::ifstream k(dir + "/grid.img.k", std::ios::binary);
        if (!k) return false;
        char magic[8];
        int hdr[6];
        double sc[4];
        k.read(magic, sizeof(magic));
        k.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
        k.read(reinterpret_cast<char*>(sc), sizeof(sc));
        const int np = cfg.pad ? (cfg.n_pad ? cfg.n_pad : 2 * cfg.N) : 0;
        if (!k || std::memcmp(magic, kSideMagic, sizeof(magic)) != 0 || hdr[0] != cfg.N || hdr[1] != cfg.nflds ||
            hdr[2] != cfg.pad || hdr[3] != np || hdr[5] != run_number || sc[0] != t0 || sc[1] != a || sc[2] != ad)
            return false;
        std::vector<double> fk(half_len() * 2, 0.0), pk(half_len() * 2, 0.0);
        k.read(reinterpret_cast<char*>(fk.data()), half_len() * cfg.nflds * sizeof(double));
        k.read(reinterpret_cast<char*>(pk.data()), half_len() * cfg.nflds * sizeof(double));
        std::vector<double> u, pu;
        if (hdr[4] && cfg.with_gw) {
            u.resize(6 * half_len());
            pu.resize(6 * half_len());
            k.read(reinterpret_cast<char*>(u.data()), u.size() * sizeof(double));
            k.read(reinterpret_cast<char*>(pu.data()), pu.size() * sizeof(double));
        }
        if (!k) return false;
        check(psc_load_ic_k(ctx, 0, fk.data(), pk.data(), a, ad), "psc_load_ic_k");
        check(psc_set_scalars(ctx, 0, t0, a, ad, sc[3]), "psc_set_scalars");
        if (!u.empty()) {
            check(psc_load_gw_k(ctx, 0, u.data(), pu.data()), "psc_load_gw_k");
            gw_restored = true;
        }
        return true;
    }
    bool gw_restored = false;

    // Reads grid.img into the context; returns false when there is none.
    bool read_grid(int& run_number, double& t0) {
        std::ifstream g(dir + "/grid.img", std::ios::binary);
        if (!g) return false;
        const size_t n3 = (size_t)cfg.N * cfg.N * cfg.N;
        std::vector<double> f(n3 * cfg.nflds), p(n3 * cfg.nflds);
        double a = 0, ad = 0;
        g.read(reinterpret_cast<char*>(&run_number), sizeof(run_number));
        g.read(reinterpret_cast<char*>(&t0), sizeof(double));
        g.read(reinterpret_cast<char*>(&a), sizeof(double));
        g.read(reinterpret_cast<char*>(&ad), sizeof(double));
        g.read(reinterpret_cast<char*>(f.data()), f.size() * sizeof(double));
        g.read(reinterpret_cast<char*>(p.data()), p.size() * sizeof(double));
        if (!g) {
            std::fprintf(stderr, "psc_run: %s/grid.img is truncated or from another lattice size\n", dir.c_str());
            std::exit(2);
        }
        if (read_side(run_number, t0, a, ad)) {
            std::printf("k-space state read from grid.img.k\n");
            return true;
        }
        check(psc_load_fields(ctx, 0, f.data(), p.data(), t0, a, ad), "psc_load_fields");
        return true;
    }

    void write_info(double dt_last, long long steps) {
        std::ofstream info(path("info"), std::ios::out);
        const double L = 2 * M_PI / cfg.kIR;
        info << "--------------------------\nModel Specific Information\n--------------------------\n";
        info << psc_model_name() << " model\n" << psc_model_description() << "\n";   // modelinfo_ (model.h:10-19)
        info << "m2 = " << model.m2 << ", g2 = " << model.g2 << ", lambda = " << model.lambda << "\n";
        info << "\n--------------------------\nGeneral Program Information\n-----------------------------\n";
        info << "Grid size=" << cfg.N << "^3\n";
        info << "Number of fields=" << cfg.nflds << "\n";
        info << "L=" << L << "\n";
        info << "kIR=" << cfg.kIR << "\n";
        info << std::setprecision(6) << "dt=" << dt_last << ", dt/dx=" << dt_last / (L / cfg.N) << "\n";
        info << (cfg.expansion == PSC_EXPANSION_STATIC ? "No expansion\n" : "Self-consistent expansion\n");
        info << "Integrator=" << (cfg.integrator == PSC_Y4 ? "Yoshida-4" : "velocity Verlet") << "\n";
        info << "Precision=" << (cfg.precision == PSC_FP32 ? "FP32 fields, FP64 background" : "FP64") << "\n";
        info << "Steps=" << steps << "\n";
        info << "Launch shape=" << psc_shape_name(ctx) << "\n";
    }
};

}  // namespace

int main(int argc, char** argv) {
    Params p;
    bool dry = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help" || a == "-h") {
            usage();
            return 0;
        }
        if (a == "--dry-run") {
            dry = true;
            continue;
        }
        if (a.find('=') != std::string::npos) {
            parse_line(a, p);
            continue;
        }
        std::ifstream in(a);
        if (!in) {
            std::fprintf(stderr, "psc_run: cannot read %s\n", a.c_str());
            return 2;
        }
        std::string line;
        while (std::getline(in, line)) parse_line(line, p);
    }
    for (auto& kv : p.kv)
        if (!known_key(kv.first)) {
            std::fprintf(stderr, "psc_run: unknown parameter %s\n", kv.first.c_str());
            return 2;
        }

    Run r;
    psc_config& c = r.cfg;
    c.N = (int)p.num("N", 64);
    c.nflds = (int)p.num("nflds", 2);
    const std::string integ = p.get("integrator", "VV");
    c.integrator = (integ == "Y4" || integ == "4") ? PSC_Y4 : PSC_VV2;
    c.yoshida_exact = (int)p.num("yoshida_exact", 0);
    c.precision = p.get("precision", "fp64") == "fp32" ? PSC_FP32 : PSC_FP64;
    c.batch = 1;
    c.with_gw = (int)p.num("with_gw", 0);
    c.pad = (int)p.num("pad", 0);
    c.n_pad = (int)p.num("N_pad", 0);
    c.alpha = (int)p.num("alpha", 0);
    c.expansion = (int)p.num("expansion", 2) == 0 ? PSC_EXPANSION_STATIC : PSC_EXPANSION_SELF;
    c.use_graph = (int)p.num("use_graph", 1);
    c.autotune = (int)p.num("autotune", 0);
    c.kIR = p.num("kIR", 25.0);
    c.fStar = p.num("fStar", 1.0);
    c.omegaStar = p.num("omegaStar", 1.0);
    c.MPl = p.num("MPl", 1.0);
    r.model.m2 = p.num("m2", 1.0);
    r.model.g2 = p.num("g2", 2.5e5);
    r.model.lambda = p.num("lambda", 0.0);
    {
        const std::vector<double> cs = p.list("couplings", {});
        for (size_t i = 0; i < cs.size() && i < 5; ++i) r.model.c[i] = cs[i];
    }
    const std::vector<double> f0 = p.list("initfield", {1.0, 0.0}), fd0 = p.list("initderivs", {-0.7, 0.0});
    for (int q = 0; q < 2; ++q) {
        r.ic.f0[q] = q < (int)f0.size() ? f0[q] : 0.0;
        r.ic.fd0[q] = q < (int)fd0.size() ? fd0[q] : 0.0;
    }
    r.ic.a0 = 1.0;
    r.ic.seed = (int)p.num("seed", 1);
    r.ic.kcutoff = p.num("kcutoff", 0.0);
    r.dir = p.get("dir", "output");
    r.gw_proj = p.get("gw_projection", "ref") == "tt" ? PSC_GW_PROJ_TT : PSC_GW_PROJ_REFERENCE;
    const std::string ic_where = p.get("ic", "host");
    if (ic_where != "host" && ic_where != "device") {
        std::fprintf(stderr, "psc_run: ic must be host or device\n");
        return 2;
    }
    const bool ic_device = ic_where == "device";
    const std::string snap_format = p.get("snap_format", "h5");
    if (snap_format != "h5" && snap_format != "raw") {
        std::fprintf(stderr, "psc_run: snap_format must be h5 or raw\n");
        return 2;
    }
    r.snap_h5 = snap_format == "h5";
    r.snap_energy = (int)p.num("snap_energy", 1) != 0;
    r.hdf5_lib = p.get("hdf5_lib", "");
    const double kmax = c.kIR * c.N * std::sqrt(3.0) / 2.0;
    auto calculate_dt = [&](double a) { return 2.0 / (kmax / a) / 10; };   // pspectcosmo.cpp:213-233
    r.auto_dt = p.get("dt", "auto") == "auto";
    c.dt = r.auto_dt ? calculate_dt(r.ic.a0) : p.num("dt", 0.005);
    // the reference's per-step dt runs on the device, inside the captured steps
    c.dt_rule = r.auto_dt ? PSC_DT_REFERENCE : PSC_DT_FIXED;
    const long long nsteps = (long long)p.num("nsteps", 200);
    const double tf = p.num("tf", r.auto_dt ? 1.0 : nsteps * c.dt);
    r.tf = tf;
    r.lattice_times = p.list("store_lattice_times", {});
    const double tOutFreq = p.num("tOutputFreq", 0.01), tOutInfreq = p.num("tOutputInfreq", 0.1);
    const double tSnap = p.num("tSnapFreq", 0.0);
    const long long freq_steps = (long long)p.num("freq_steps", 20), infreq_steps = (long long)p.num("infreq_steps", 20);
    const long long snap_steps = (long long)p.num("snap_steps", 0);
    r.nbins = psc_num_bins(c.N);

    if (dry) {
        std::printf("N=%d nflds=%d integrator=%s precision=%s expansion=%d alpha=%d kIR=%g dt=%s%g "
                    "m2=%g g2=%g initfield=%g,%g initderivs=%g,%g seed=%d with_gw=%d pad=%d dir=%s\n",
                    c.N, c.nflds, c.integrator == PSC_Y4 ? "Y4" : "VV", c.precision == PSC_FP32 ? "fp32" : "fp64",
                    c.expansion, c.alpha, c.kIR, r.auto_dt ? "auto:" : "", c.dt, r.model.m2, r.model.g2,
                    r.ic.f0[0], r.ic.f0[1], r.ic.fd0[0], r.ic.fd0[1], r.ic.seed, c.with_gw, c.pad, r.dir.c_str());
        if (r.auto_dt)
            std::printf("tf=%g tOutputFreq=%g tOutputInfreq=%g tSnapFreq=%g\n", tf, tOutFreq, tOutInfreq, tSnap);
        else
            std::printf("nsteps=%lld freq_steps=%lld infreq_steps=%lld snap_steps=%lld\n", nsteps, freq_steps,
                        infreq_steps, snap_steps);
        return 0;
    }

    if (mkdir(r.dir.c_str(), 0755) != 0 && errno != EEXIST) {
        std::fprintf(stderr, "psc_run: cannot create %s\n", r.dir.c_str());
        return 2;
    }
    // initial conditions (initialize(), initialize.cpp:624-800)
    const size_t nk = (size_t)c.nflds * c.N * c.N * (c.N / 2 + 1) * 2;   // FFTW half spectra
    std::vector<double> fk, pk;
    double sc0[4];
    if (!ic_device) {
        fk.resize(nk);
        pk.resize(nk);
        r.check(psc_generate_ic(&c, &r.model, &r.ic, fk.data(), pk.data(), sc0), "psc_generate_ic");
    }
    r.check(psc_create(&c, &r.ctx), "psc_create");
    r.check(psc_set_model(r.ctx, 0, &r.model), "psc_set_model");
    const bool checkpointing = (int)p.num("checkpoint", 0) != 0;
    int run_number = 0;
    double t_start = 0.0;
    bool resumed = false;
    if ((int)p.num("continue_run", 0) != 0 && r.read_grid(run_number, t_start)) {
        ++run_number;
        resumed = true;
        std::printf("Data read. Resuming run at t = %g\n", t_start);
        if (c.with_gw && !r.gw_restored)
            std::fprintf(stderr, "psc_run: grid.img holds no GW state; GW restart from vacuum\n");
    } else if (ic_device) {
        r.check(psc_generate_ic_device(r.ctx, 0, &r.ic, sc0), "psc_generate_ic_device");
    } else {
        r.check(psc_load_ic_k(r.ctx, 0, fk.data(), pk.data(), sc0[0], sc0[1]), "psc_load_ic_k");
    }
    if (c.with_gw && r.gw_restored) {
        // the GW state came with the k-space image
    } else if (c.with_gw && ic_device) {
        r.check(psc_generate_gw_ic_device(r.ctx, 0, &r.ic), "psc_generate_gw_ic_device");
    } else if (c.with_gw) {
        const size_t gk = (size_t)6 * c.N * c.N * (c.N / 2 + 1) * 2;
        std::vector<double> u(gk), pu(gk);
        r.check(psc_generate_gw_ic(&c, &r.model, &r.ic, u.data(), pu.data()), "psc_generate_gw_ic");
        r.check(psc_load_gw_k(r.ctx, 0, u.data(), pu.data()), "psc_load_gw_k");
    }
    if (c.autotune) {
        char rep[1024];
        r.check(psc_autotune(r.ctx, 2, rep, sizeof(rep)), "psc_autotune");
        std::printf("autotune: %s -> %s\n", rep, psc_shape_name(r.ctx));
    }
    r.open_outputs();

    long long steps = 0;
    int snaps = 0;
    double dt_last = c.dt;
    if (!r.auto_dt) {
        // fixed dt with a step-count cadence
        long long done = resumed ? std::llround(t_start / c.dt) : 0;
        if (done >= nsteps) {
            std::fprintf(stderr, "psc_run: the lattice image is at step %lld of %lld; raise nsteps to continue\n",
                         done, nsteps);
            return 1;
        }
        while (done < nsteps) {
            long long next = nsteps;
            auto upto = [&](long long every) {
                if (every > 0) next = std::min(next, (done / every + 1) * every);
            };
            upto(freq_steps);
            upto(infreq_steps);
            upto(snap_steps);
            r.check(psc_step(r.ctx, (int)(next - done)), "psc_step");
            done = next;
            if (freq_steps > 0 && done % freq_steps == 0) r.save_freq();
            if (infreq_steps > 0 && done % infreq_steps == 0) {
                r.save_infreq();
                if (checkpointing) r.write_grid(run_number);
                std::printf("Lattice Step %lld done. Current time: %g\n", done, done * c.dt);
                std::fflush(stdout);
            }
            if (snap_steps > 0 && done % snap_steps == 0) r.save_snap(++snaps);
        }
        steps = done;
    } else {
        // the reference's loop (pspectcosmo.cpp:601-697)
        double t = t_start, a = r.ic.a0;
        if (resumed) {
            double s0[4];
            r.check(psc_get_scalars(r.ctx, 0, s0), "psc_get_scalars");
            a = s0[1];
            if (t >= tf) {
                std::fprintf(stderr, "psc_run: the lattice image is at t = %g >= tf; raise tf to continue\n", t);
                return 1;
            }
        }
        // next output times after t (pspectcosmo.cpp:366-367 for a new run)
        auto next_after = [&](double every) { return every * (std::floor(t / every + 1e-9) + 1); };
        double next_freq = next_after(tOutFreq), next_infreq = next_after(tOutInfreq);
        double next_snap = tSnap > 0 ? next_after(tSnap) : 1e300;
        // Each call runs the steps up to the next output time: the device takes
        // the reference's dt per step, capped at tOutputFreq and clamped to
        // land on the output time, after which its steps are empty.  The
        // number of steps of a call is estimated from the current dt (a grows,
        // so dt does not shrink); a call that falls short is followed by
        // another.
        double target = -1.0;
        while (t < tf) {
            const double t_out = std::min(next_freq, next_infreq);
            if (t_out != target) {
                r.check(psc_set_output_times(r.ctx, t_out, tOutFreq), "psc_set_output_times");
                target = t_out;
            }
            double dt = std::min(calculate_dt(a), tOutFreq);
            if (t + dt > t_out) dt = t_out - t;
            const double n = std::ceil((t_out - t) / dt - 1e-9);
            // the reference stops at the first step that reaches tf: near tf
            // the steps go one per call
            const double n_tf = std::ceil((tf - t) / dt - 1e-9);
            const int calls = (int)std::max(1.0, std::min(n_tf <= n ? n_tf - 1 : n, 1e6));
            r.check(psc_step(r.ctx, calls), "psc_step");
            double s[4];
            r.check(psc_get_scalars(r.ctx, 0, s), "psc_get_scalars");
            t = s[0];
            a = s[1];
            long long done_steps = 0;
            r.check(psc_get_step_count(r.ctx, 0, &done_steps), "psc_get_step_count");
            steps = done_steps;
            dt_last = dt;
            if (std::fabs(t - next_freq) < 1e-10) {
                next_freq += tOutFreq;
                r.save_freq();
            }
            if (std::fabs(t - next_infreq) < 1e-10) {
                next_infreq += tOutInfreq;
                r.save_infreq();
                if (checkpointing) r.write_grid(run_number);
                std::printf("Lattice Step %lld done. Current time: %g\n", steps, t);
                std::fflush(stdout);
            }
            if (std::fabs(t - next_snap) < 1e-10) {
                next_snap += tSnap;
                r.save_snap(++snaps);
            }
        }
    }
    if (checkpointing) r.write_grid(run_number);
    r.write_info(dt_last, steps);
    psc_destroy(r.ctx);
    std::printf("PSpectCosmo MI355X run finished: %lld steps\n", steps);
    return 0;
}
