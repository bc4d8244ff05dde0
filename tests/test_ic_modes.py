"""The visiting order and stream positions of the device IC generator
(pspectcosmo_amd/csrc/psc_ic_modes.hpp), on the CPU: a small program built
from the header lists mode_at(t) for every work item and Stream draws at
chosen positions, checked against

* the loop nest of initialize() (initialize.cpp:685-775) written out here:
  every work item lands on the mode with the same position in the visiting
  order, conjugates included, and the work items cover every mode once;
* the full spectrum: each index is written exactly once (mode or conjugate);
* the oracle's sequential Park-Miller stream (pinned by its known-answer
  test), including the seed < 1 debug constant."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r'''
#include "psc_ic_modes.hpp"
#include <cstdio>
#include <cstdlib>
#include <string>
int main(int argc, char** argv) {
    const int N = std::atoi(argv[1]);
    const bool dc = std::atoi(argv[2]) != 0;
    if (argc > 3 && std::string(argv[3]) == "rank") {   // rank mode: kIR, kcutoff
        psc_config c{};
        c.N = N;
        c.kIR = std::atof(argv[4]);
        const psc_ic::Cut cut = psc_ic::make_cut(c, std::atof(argv[5]));
        psc_ic::KeptIndex kept;
        std::vector<uint64_t> rowpre;
        const unsigned long long total = psc_ic::build_kept_index(N, dc, cut, kept, rowpre);
        std::printf("%lld %llu\n", (long long)cut.smax, total);
        const unsigned long long n = psc_ic::mode_count(N, dc);
        for (unsigned long long t = 0; t < n; ++t) {
            const psc_ic::Mode m = psc_ic::mode_at(N, dc, t);
            const long long px = m.i <= N / 2 ? m.i : m.i - N, py = m.j <= N / 2 ? m.j : m.j - N;
            std::printf("%llu %llu %d\n", m.v, (unsigned long long)kept.rank(N, dc, m, t),
                        cut.keeps(px * px + py * py + (long long)m.k * m.k) ? 1 : 0);
        }
        return 0;
    }
    if (argc > 3) {   // stream mode: seed, then draw positions
        const int seed = std::atoi(argv[3]);
        for (int a = 4; a < argc; ++a) {
            const unsigned long long d = std::strtoull(argv[a], nullptr, 10);
            psc_ic::Stream s = psc_ic::Stream::at(seed, d);
            std::printf("%.17g", s.uniform());
            s.next();
            std::printf(" %.17g\n", s.uniform());
        }
        return 0;
    }
    const unsigned long long n = psc_ic::mode_count(N, dc);
    for (unsigned long long t = 0; t < n; ++t) {
        const psc_ic::Mode m = psc_ic::mode_at(N, dc, t);
        std::printf("%llu %d %d %d %d %d %d\n", m.v, m.i, m.j, m.k, m.ci, m.cj, m.real ? 1 : 0);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    d = tmp_path_factory.mktemp("icmodes")
    src = d / "m.cpp"
    src.write_text(PROG)
    exe = d / "m"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "pspectcosmo_amd", "csrc"),
                           str(src), "-o", str(exe)])
    return str(exe)


def reference_order(N, dc):
    """(i, j, k, ci, cj, real) in the order of initialize.cpp:685-775."""
    h = N // 2
    out = []
    for k in (0, h):
        for i in (0, h):
            for j in range(1, h):
                out.append((i, j, k, i, N - j, 0))
        for j in (0, h):
            for i in range(1, h):
                out.append((i, j, k, N - i, j, 0))
        for i in range(1, h):
            for j in range(1, h):
                out.append((i, j, k, N - i, N - j, 0))
            for j in range(h + 1, N):
                out.append((i, j, k, N - i, N - j, 0))
    for k in range(1, h):
        for i in range(N):
            for j in range(N):
                out.append((i, j, k, -1, -1, 0))
    for k in (0, h):
        for i in (0, h):
            for j in (0, h):
                if not dc and i == j == k == 0:
                    continue
                out.append((i, j, k, -1, -1, 1))
    return out


@pytest.mark.parametrize("N", [16, 32])
@pytest.mark.parametrize("dc", [1, 0])
def test_mode_order_and_coverage(prog, N, dc):
    rows = np.loadtxt(subprocess.check_output([prog, str(N), str(dc)], text=True).splitlines(), dtype=np.int64)
    ref = reference_order(N, dc)
    assert len(rows) == len(ref)
    v = rows[:, 0]
    assert np.array_equal(np.sort(v), np.arange(len(ref)))          # a bijection onto the order
    for r in rows:
        assert tuple(int(x) for x in r[1:]) == ref[r[0]]
    # every full-spectrum index once: the mode, and its negative unless real
    seen = np.zeros((N, N, N), dtype=np.int64)
    for _, i, j, k, ci, cj, real in rows:
        seen[i, j, k] += 1
        if not real:
            seen[(N - i) % N, (N - j) % N, (N - k) % N] += 1
            if ci >= 0:   # the in-plane conjugate of the reference is that negative
                assert (ci, cj) == ((N - i) % N, (N - j) % N) and k in (0, N // 2)
    expect = np.ones((N, N, N), dtype=np.int64)
    if not dc:
        expect[0, 0, 0] = 0
    assert np.array_equal(seen, expect)


def test_stream_positions_match_the_sequential_stream(prog):
    seed = 12345
    seq = O.rng_draws(seed, 5000)
    pos = [0, 1, 2, 3, 4, 97, 998, 4998]
    out = subprocess.check_output([prog, "16", "1", str(seed), *map(str, pos)], text=True).split()
    got = np.array(out, dtype=float).reshape(-1, 2)
    for (a, b), d in zip(got, pos):
        assert a == seq[d] and b == seq[d + 1]
    # far positions: 16807^(100 + d) by repeated squaring against stepping
    far = subprocess.check_output([prog, "16", "1", "7", "1000000"], text=True).split()
    seq7 = O.rng_draws(7, 1000002)
    assert float(far[0]) == seq7[1000000] and float(far[1]) == seq7[1000001]
    dbg = subprocess.check_output([prog, "16", "1", "0", "5"], text=True).split()
    assert float(dbg[0]) == 0.33 and float(dbg[1]) == 0.33


@pytest.mark.parametrize("N,dc,kIR,kcutoff", [(16, 1, 1.0, 5.0), (16, 0, 1.0, 5.0), (32, 1, 0.5, 4.3),
                                              (32, 0, 0.5, 4.3), (16, 1, 1.0, 0.0), (16, 1, 1.0, 100.0),
                                              (16, 1, 1.0, 0.5)])
def test_kept_rank_counts_the_kept_modes_before_each(prog, N, dc, kIR, kcutoff):
    """With the kcutoff low-pass every kept mode's draw offset / 4 is the
    number of kept modes before it in the visiting order (initialize.cpp:
    138-145 returns before drawing), here counted by brute force over the loop
    nest, with the cut decided in double exactly as the reference does:
    kIR^2 (px^2 + py^2 + pz^2) > kcutoff^2, for kcutoff < 2 pi N / L."""
    out = subprocess.check_output([prog, str(N), str(dc), "rank", repr(kIR), repr(kcutoff)], text=True).splitlines()
    smax, total = (int(x) for x in out[0].split())
    rows = np.loadtxt(out[1:], dtype=np.int64)
    ref = reference_order(N, dc)
    L = 2.0 * np.pi / kIR
    k2cut = kcutoff * kcutoff if kcutoff < 2.0 * np.pi * N / L else 0.0
    sidx = lambda i: i if i <= N // 2 else i - N
    keep = []
    for (i, j, k, *_rest) in ref:
        p2 = (kIR * kIR) * float(sidx(i) ** 2 + sidx(j) ** 2 + k * k)
        keep.append(not (k2cut > 0 and p2 > k2cut))
    keep = np.array(keep)
    before = np.concatenate([[0], np.cumsum(keep)[:-1]])
    v = rows[:, 0]
    assert np.array_equal(rows[:, 2].astype(bool), keep[v])
    kept_rows = rows[rows[:, 2] == 1]
    assert np.array_equal(kept_rows[:, 1], before[kept_rows[:, 0]])
    assert total == int(keep.sum())
    if k2cut == 0:
        assert smax == -1 and total == len(ref)
    else:
        assert 0 < total < len(ref) or kcutoff < kIR
