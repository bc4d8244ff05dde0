"""The command-line driver psc_run (host C++ over the C ABI): parameter
parsing without touching a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PSC_RUN = os.path.join(ROOT, "pspectcosmo_amd", "psc_run")

pytestmark = pytest.mark.skipif(not os.path.exists(PSC_RUN), reason="psc_run not built")


def run(*args):
    return subprocess.run([PSC_RUN, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=60)


def test_help_lists_the_parameters():
    r = run("--help")
    assert r.returncode == 0
    for key in ("N", "kIR", "dt", "g2", "initfield", "tOutputFreq", "freq_steps", "with_gw"):
        assert f"  {key} " in r.stdout


def test_parameter_file_and_overrides(tmp_path):
    pf = tmp_path / "params.txt"
    pf.write_text("# config 0 of the benchmark\nN = 64\nkIR = 25\ndt = 0.005\nnsteps = 200\n"
                  "g2 = 2.5e5\ninitfield = 1.0, 0\ninitderivs = -0.7, 0\nfreq_steps = 20\n")
    r = run(str(pf), "precision=fp32", "--dry-run")
    assert r.returncode == 0, r.stderr
    assert "N=64" in r.stdout and "dt=0.005" in r.stdout and "precision=fp32" in r.stdout
    assert "initfield=1,0 initderivs=-0.7,0" in r.stdout and "nsteps=200 freq_steps=20" in r.stdout


def test_auto_dt_is_the_default(tmp_path):
    r = run("N=64", "kIR=25", "--dry-run")
    assert r.returncode == 0
    # calculate_dt (pspectcosmo.cpp:213-233) at a = 1: 2 / (25 sqrt(3) 32) / 10
    assert "dt=auto:0.000144338" in r.stdout


def test_unknown_parameter_is_an_error():
    r = run("Nx=64", "--dry-run")
    assert r.returncode == 2 and "unknown parameter Nx" in r.stderr


def test_snapshot_format_is_checked():
    r = run("snap_format=netcdf", "--dry-run")
    assert r.returncode == 2 and "snap_format" in r.stderr
    for fmt in ("h5", "raw"):
        assert run(f"snap_format={fmt}", "--dry-run").returncode == 0


def test_ic_source_is_checked():
    r = run("ic=gpu", "--dry-run")
    assert r.returncode == 2 and "ic must be" in r.stderr
    assert run("ic=device", "--dry-run").returncode == 0
