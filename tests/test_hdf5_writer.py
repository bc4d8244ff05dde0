"""The snapshot writer (pspectcosmo_amd/csrc/psc_hdf5.hpp) on the CPU: a small
program built from the header writes two N^3 datasets, which the HDF5 C
library reads back bit for bit with their names, rank, shape and native
double type (the layout of snapshots, output.cpp:1541-1648)."""
import os
import subprocess

import numpy as np
import pytest

from tests import _h5

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r'''
#include "psc_hdf5.hpp"
#include <cstdio>
#include <cmath>
int main(int argc, char** argv) {
    const int N = 6;
    std::vector<double> a(N * N * N), b(N * N * N);
    for (int i = 0; i < N * N * N; ++i) { a[i] = i + 0.25; b[i] = std::sin(0.1 * i) * 1e-300; }
    psc_io::Hdf5 h;
    std::string err = h.load(argc > 2 ? argv[2] : "");
    if (err.empty()) err = h.write(argv[1], N, {{"/A", a.data()}, {"/B", b.data()}});
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    return 0;
}
'''


@pytest.mark.skipif(_h5.find_lib() is None, reason="no HDF5 library in this image")
def test_writer_round_trip(tmp_path):
    src = tmp_path / "w.cpp"
    src.write_text(PROG)
    exe = tmp_path / "w"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "pspectcosmo_amd", "csrc"),
                           str(src), "-o", str(exe), "-ldl"])
    out = tmp_path / "x.h5"
    subprocess.check_call([str(exe), str(out)])
    d = _h5.read(str(out))
    assert sorted(d) == ["/A", "/B"]
    N = 6
    i = np.arange(N ** 3)
    assert np.array_equal(d["/A"], (i + 0.25).reshape(N, N, N))
    assert np.array_equal(d["/B"], (np.sin(0.1 * i) * 1e-300).reshape(N, N, N))
    # an unusable hdf5_lib path falls back to the loader's search and the default path
    r = subprocess.run([str(exe), str(tmp_path / "y.h5"), "/nonexistent/libhdf5.so"], stderr=subprocess.PIPE,
                       text=True)
    assert r.returncode == 0, r.stderr
