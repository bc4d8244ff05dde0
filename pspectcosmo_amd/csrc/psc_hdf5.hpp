// psc_hdf5.hpp -- the HDF5 writer of the snapshots (snapshots,
// output.cpp:1541-1648), shared by the driver psc_run and its tests.
#pragma once

#include <cstdint>
#include <dlfcn.h>
#include <string>
#include <utility>
#include <vector>

namespace psc_io {

// The few HDF5 C-library calls the snapshots need, resolved with dlopen so the
// driver has no link-time dependency on HDF5.  Types and constants are those
// of the 1.10 ABI (hid_t 64-bit), which load() checks against the library's
// own version: H5F_ACC_TRUNC = 2, H5P_DEFAULT = H5S_ALL = 0.
struct Hdf5 {
    using hid = int64_t;
    using herr = int;
    using hsize = unsigned long long;
    void* so = nullptr;
    herr (*open)() = nullptr;
    herr (*version)(unsigned*, unsigned*, unsigned*) = nullptr;
    hid (*fcreate)(const char*, unsigned, hid, hid) = nullptr;
    hid (*screate_simple)(int, const hsize*, const hsize*) = nullptr;
    hid (*dcreate2)(hid, const char*, hid, hid, hid, hid, hid) = nullptr;
    herr (*dwrite)(hid, hid, hid, hid, hid, const void*) = nullptr;
    herr (*dclose)(hid) = nullptr;
    herr (*sclose)(hid) = nullptr;
    herr (*fclose)(hid) = nullptr;
    const hid* native_double = nullptr;

    template <class F>
    bool sym(F& f, const char* name) {
        f = reinterpret_cast<F>(dlsym(so, name));
        return f != nullptr;
    }
    // Empty string on success, else what went wrong.
    std::string load(const std::string& lib) {
        if (so) return "";
        std::vector<std::string> names;
        if (!lib.empty()) names.push_back(lib);
        names.push_back("libhdf5.so");
        names.push_back("/opt/conda/lib/libhdf5.so");
        std::string tried;
        for (auto& n : names) {
            so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (so) break;
            tried += " " + n;
        }
        if (!so) return "no HDF5 library found (tried" + tried + "); set hdf5_lib= or snap_format=raw";
        bool ok = sym(open, "H5open") && sym(version, "H5get_libversion") && sym(fcreate, "H5Fcreate") &&
                  sym(screate_simple, "H5Screate_simple") && sym(dcreate2, "H5Dcreate2") &&
                  sym(dwrite, "H5Dwrite") && sym(dclose, "H5Dclose") && sym(sclose, "H5Sclose") &&
                  sym(fclose, "H5Fclose");
        native_double = static_cast<const hid*>(dlsym(so, "H5T_NATIVE_DOUBLE_g"));
        if (!ok || !native_double) return "the HDF5 library lacks an expected symbol";
        unsigned maj = 0, min = 0, rel = 0;
        if (open() < 0 || version(&maj, &min, &rel) < 0) return "H5open failed";
        if (maj != 1 || min < 10) return "HDF5 " + std::to_string(maj) + "." + std::to_string(min) + " found; 1.10 or later needed";
        return "";
    }
    // One file of N^3 double datasets (H5F_ACC_TRUNC, default property lists).
    std::string write(const std::string& file, int N, const std::vector<std::pair<std::string, const double*>>& sets) {
        const hid f = fcreate(file.c_str(), 2u, 0, 0);
        if (f < 0) return "cannot create " + file;
        const hsize dims[3] = {(hsize)N, (hsize)N, (hsize)N};
        const hid space = screate_simple(3, dims, nullptr);
        std::string err;
        if (space < 0) err = "H5Screate_simple failed";
        for (size_t i = 0; err.empty() && i < sets.size(); ++i) {
            const hid d = dcreate2(f, sets[i].first.c_str(), *native_double, space, 0, 0, 0);
            if (d < 0) {
                err = "cannot create dataset " + sets[i].first;
                break;
            }
            if (dwrite(d, *native_double, 0, 0, 0, sets[i].second) < 0) err = "cannot write dataset " + sets[i].first;
            dclose(d);
        }
        if (space >= 0) sclose(space);
        if (fclose(f) < 0 && err.empty()) err = "cannot close " + file;
        return err;
    }
};

}  // namespace psc_io
