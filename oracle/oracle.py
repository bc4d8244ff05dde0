"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU FP64 oracle
(oracle/psc_oracle.cpp).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; the product library never
does."""
import ctypes as C
import os
import subprocess

import numpy as np

from pspectcosmo_amd.abi import (PscConfig, PscModel, PscIcParams, PSC_ENERGY_LEN,
                                 PSC_SPECTRA_COLS, PSC_GW_SPECTRA_COLS, num_bins, half_len)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
DEFAULT_MODEL = "preheat"
_libs = {}


def lib_path(model_name=None):
    """liboracle.so is built with the default model header (preheat),
    liboracle_<name>.so with another built-in model's header."""
    if not model_name or model_name == DEFAULT_MODEL:
        return _LIB_PATH
    return os.path.join(_HERE, f"liboracle_{model_name}.so")

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force=False):
    if force or not all(os.path.exists(lib_path(m)) for m in (None, "lambda_phi4")):
        subprocess.check_call(["make", "-C", _HERE] + (["-B"] if force else []) + ["all"],
                              stdout=subprocess.DEVNULL)


def lib(model_name=None):
    path = lib_path(model_name)
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_park_miller_iterate.restype = C.c_longlong
        L.orc_park_miller_iterate.argtypes = [C.c_longlong, C.c_longlong]
        L.orc_rng_draws.argtypes = [C.c_int, C.c_int, _D]
        L.orc_fft_r2c_3d.argtypes = [C.c_int, _D, _D]
        L.orc_fft_c2r_3d.argtypes = [C.c_int, _D, _D]
        L.orc_fft_1d.argtypes = [C.c_int, _D, C.c_int]
        L.orc_generate_ic.argtypes = [P(PscConfig), P(PscModel), P(PscIcParams), _D, _D, _D]
        L.orc_generate_gw_ic.argtypes = [P(PscConfig), P(PscModel), P(PscIcParams), _D, _D]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P(PscConfig), P(PscModel)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_load.argtypes = [C.c_void_p, _D, _D, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_load_gw.argtypes = [C.c_void_p, _D, _D]
        L.orc_get_gw.argtypes = [C.c_void_p, _D, _D]
        L.orc_kernel_a.restype = C.c_double
        L.orc_kernel_a.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.orc_plant_gradient_power.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_k.argtypes = [C.c_void_p, _D, _D]
        L.orc_get_scalars.argtypes = [C.c_void_p, _D]
        L.orc_get_fields.argtypes = [C.c_void_p, _D, _D]
        L.orc_field_derivative.argtypes = [C.c_void_p, C.c_int, C.c_int, _D]
        L.orc_vprime_k.argtypes = [C.c_void_p, _D]
        L.orc_energies.argtypes = [C.c_void_p, _D]
        L.orc_energy_densities.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.orc_spectra.argtypes = [C.c_void_p, C.c_int, _D, C.c_int]
        L.orc_gw_spectrum.argtypes = [C.c_void_p, C.c_int, _D, C.c_int, _D]
        L.orc_gw_omega_log.argtypes = [C.c_void_p, C.c_int, C.c_int, _D, _D]
        L.orc_adjoint_chi2.argtypes = [P(PscConfig), P(PscModel), _D, _D, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_int, P(C.c_double), P(C.c_double),
                                       P(C.c_double), C.c_void_p]
        L.orc_adjoint_gw.argtypes = [P(PscConfig), P(PscModel), _D, _D, _D, _D, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, P(C.c_double),
                                     P(C.c_double), P(C.c_double), C.c_void_p]
        L.orc_set_fft_backend.argtypes = [C.c_int, C.c_int]
        L.orc_fft_backend_name.restype = C.c_char_p
        L.orc_model_name.restype = C.c_char_p
        L.orc_model_eval.argtypes = [P(PscModel), C.c_int, _D, _D, C.c_int, _D]
        _libs[path] = L
    return _libs[path]


def model_name(model_name=None):
    return lib(model_name).orc_model_name().decode()


def model_eval(model, nflds, f, df=(0.0, 0.0), coupling=0, model_name=None):
    """The model header at one point: dict of terms, dvdf, d2Vdf2, meff,
    delta (FP64), delta32 (FP32), dV_dc, d2V_dfdc, nterms; ok = coupling
    supported."""
    out = np.zeros(19)
    rc = lib(model_name).orc_model_eval(C.byref(model), int(nflds), np.asarray(f, dtype=np.float64),
                                        np.asarray(df, dtype=np.float64), int(coupling), out)
    return {"terms": out[0:4], "dvdf": out[4:6], "d2Vdf2": out[6:9], "meff": out[9:11],
            "delta": out[11:13], "delta32": out[13:15], "dV_dc": out[15], "d2V_dfdc": out[16:18],
            "nterms": int(out[18]), "ok": rc == 0}


def set_threads(n):
    return lib().orc_set_threads(int(n))


def set_fft_backend(backend, threads=1):
    """0: the oracle's radix-2 FFT; 1: the FFTW3 API (libfftw3 or MKL) with
    `threads` threads.  Returns the backend's name, or None if it is absent."""
    rc = lib().orc_set_fft_backend(int(backend), int(threads))
    return lib().orc_fft_backend_name().decode() if rc == 0 else None


def park_miller_iterate(state, steps):
    return lib().orc_park_miller_iterate(state, steps)


def rng_draws(seed, n):
    out = np.zeros(n)
    lib().orc_rng_draws(seed, n, out)
    return out


def fft_r2c_3d(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    N = x.shape[0]
    out = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
    lib().orc_fft_r2c_3d(N, x.ravel(), out.view(np.float64).ravel())
    return out


def fft_c2r_3d(xk):
    xk = np.ascontiguousarray(xk, dtype=np.complex128)
    N = xk.shape[0]
    out = np.zeros((N, N, N))
    lib().orc_fft_c2r_3d(N, xk.view(np.float64).ravel(), out.ravel())
    return out


def fft_1d(x, sign=-1):
    x = np.array(x, dtype=np.complex128)
    buf = x.view(np.float64).copy()
    lib().orc_fft_1d(len(x), buf, sign)
    return buf.view(np.complex128)


def generate_ic(cfg, model, ic, model_name=None):
    """Half spectra (nflds, N, N, N/2+1) complex, and scalars
    (a, ad, add_printed, H, add_kernel)."""
    N, nf = cfg.N, cfg.nflds
    fk = np.zeros((nf, N, N, N // 2 + 1), dtype=np.complex128)
    pk = np.zeros_like(fk)
    sc = np.zeros(5)
    rc = lib(model_name).orc_generate_ic(C.byref(cfg), C.byref(model), C.byref(ic),
                               fk.view(np.float64).ravel(), pk.view(np.float64).ravel(), sc)
    assert rc == 0, rc
    return fk, pk, sc


def generate_gw_ic(cfg, model, ic, model_name=None):
    N = cfg.N
    u = np.zeros((6, N, N, N // 2 + 1), dtype=np.complex128)
    pu = np.zeros_like(u)
    rc = lib(model_name).orc_generate_gw_ic(C.byref(cfg), C.byref(model), C.byref(ic),
                                  u.view(np.float64).ravel(), pu.view(np.float64).ravel())
    assert rc == 0, rc
    return u, pu


def adjoint_chi2(cfg, model, fk, pk, a0, phi0, dphi0, chi0, nsteps, want_field=True, coupling=0,
                 model_name=None):
    """FP64 reference gradient of the final <chi^2>: (J, dJ/dg2, dJ/dphi0,
    dJ/dchi(x) as an (N, N, N) array or None)."""
    fk = np.ascontiguousarray(fk, dtype=np.complex128)
    pk = np.ascontiguousarray(pk, dtype=np.complex128)
    J, g2b, p0b = C.c_double(), C.c_double(), C.c_double()
    N = cfg.N
    chib = np.zeros((N, N, N)) if want_field else None
    rc = lib(model_name).orc_adjoint_chi2(C.byref(cfg), C.byref(model), fk.view(np.float64).ravel(),
                                          pk.view(np.float64).ravel(), a0, phi0, dphi0, chi0, int(nsteps),
                                          int(coupling),
                                C.byref(J), C.byref(g2b), C.byref(p0b),
                                chib.ctypes.data_as(C.c_void_p) if want_field else None)
    assert rc == 0, rc
    return J.value, g2b.value, p0b.value, chib


def adjoint_gw(cfg, model, fk, pk, u, pu, a0, phi0, dphi0, chi0, nsteps, projection, want_field=True,
               coupling=0, model_name=None):
    """FP64 reference gradient of the final GW energy fraction rho_GW / rho:
    (J, dJ/dg2, dJ/dphi0, dJ/dchi(x) or None)."""
    arrs = [np.ascontiguousarray(x, dtype=np.complex128).view(np.float64).ravel() for x in (fk, pk, u, pu)]
    J, g2b, p0b = C.c_double(), C.c_double(), C.c_double()
    N = cfg.N
    chib = np.zeros((N, N, N)) if want_field else None
    rc = lib(model_name).orc_adjoint_gw(C.byref(cfg), C.byref(model), *arrs, a0, phi0, dphi0, chi0,
                                        int(nsteps), int(projection), int(coupling), C.byref(J), C.byref(g2b), C.byref(p0b),
                              chib.ctypes.data_as(C.c_void_p) if want_field else None)
    assert rc == 0, rc
    return J.value, g2b.value, p0b.value, chib


class Oracle:
    """One lattice evolved by the FP64 CPU restatement."""

    def __init__(self, cfg, model, model_name=None):
        self.cfg, self.model = cfg, model
        self._L = lib(model_name)
        self.N, self.nf = cfg.N, cfg.nflds
        self._h = self._L.orc_create(C.byref(cfg), C.byref(model))
        if not self._h:
            raise ValueError("oracle rejected the configuration")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_destroy(h)
            self._h = None

    def load(self, fk, pk, t, a, ad, add):
        fk = np.ascontiguousarray(fk, dtype=np.complex128)
        pk = np.ascontiguousarray(pk, dtype=np.complex128)
        self._L.orc_load(self._h, fk.view(np.float64).ravel(), pk.view(np.float64).ravel(),
                       t, a, ad, add)

    def load_gw(self, u, pu):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        pu = np.ascontiguousarray(pu, dtype=np.complex128)
        assert self._L.orc_load_gw(self._h, u.view(np.float64).ravel(),
                                 pu.view(np.float64).ravel()) == 0

    def get_gw(self):
        N = self.N
        u = np.zeros((6, N, N, N // 2 + 1), dtype=np.complex128)
        pu = np.zeros_like(u)
        assert self._L.orc_get_gw(self._h, u.view(np.float64).ravel(),
                                pu.view(np.float64).ravel()) == 0
        return u, pu

    def kernel_a(self):
        return self._L.orc_kernel_a(self._h)

    def step(self, n=1):
        self._L.orc_step(self._h, int(n))

    def set_dt(self, dt):
        self._L.orc_set_dt(self._h, float(dt))

    def plant_gradient_power(self, offset):
        """Plant a wrong power of a on the gradient force (tests that must
        detect such a mistake)."""
        self._L.orc_plant_gradient_power(self._h, int(offset))

    def get_k(self):
        N = self.N
        fk = np.zeros((self.nf, N, N, N // 2 + 1), dtype=np.complex128)
        pk = np.zeros_like(fk)
        self._L.orc_get_k(self._h, fk.view(np.float64).ravel(), pk.view(np.float64).ravel())
        return fk, pk

    def scalars(self):
        s = np.zeros(4)
        self._L.orc_get_scalars(self._h, s)
        return s

    def fields(self):
        N = self.N
        f = np.zeros((self.nf, N, N, N))
        p = np.zeros_like(f)
        self._L.orc_get_fields(self._h, f.ravel(), p.ravel())
        return f, p

    def vprime_k(self):
        """V'_k half spectra per field, as the stepper's force forms them."""
        N = self.N
        out = np.zeros((self.nf, N, N, N // 2 + 1), dtype=np.complex128)
        assert self._L.orc_vprime_k(self._h, out.view(np.float64).ravel()) == 0
        return out

    def derivative(self, fld, axis):
        N = self.N
        out = np.zeros((N, N, N))
        self._L.orc_field_derivative(self._h, fld, axis, out.ravel())
        return out

    def energies(self):
        e = np.zeros(PSC_ENERGY_LEN)
        self._L.orc_energies(self._h, e)
        return e

    def energy_densities(self):
        """(pot, kin, grad, tot), each (N, N, N)."""
        N = self.N
        out = [np.zeros((N, N, N)) for _ in range(4)]
        assert self._L.orc_energy_densities(self._h, *[o.ravel() for o in out]) == 0
        return tuple(out)

    def spectra(self, fld):
        nb = num_bins(self.N)
        out = np.zeros((nb, PSC_SPECTRA_COLS))
        assert self._L.orc_spectra(self._h, fld, out.ravel(), nb) == 0
        return out

    def gw_spectrum(self, projection):
        nb = num_bins(self.N)
        out = np.zeros((nb, PSC_GW_SPECTRA_COLS))
        tot = np.zeros(2)
        assert self._L.orc_gw_spectrum(self._h, projection, out.ravel(), nb, tot) == 0
        return out, tot

    def gw_omega_log(self, projection, nlog):
        kc = np.zeros(nlog)
        om = np.zeros(nlog)
        assert self._L.orc_gw_omega_log(self._h, projection, nlog, kc, om) == 0
        return kc, om

__all__ = ["Oracle", "generate_ic", "generate_gw_ic", "fft_r2c_3d", "fft_c2r_3d", "fft_1d",
           "rng_draws", "park_miller_iterate", "set_threads", "half_len"]
