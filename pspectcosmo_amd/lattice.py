"""Python binding of libpsc.so (include/psc.h) over ctypes.

The HIP library is the only compute path: importing this module raises if the
built library is missing, and there is no CPU fallback.  Arrays crossing the
boundary are numpy FP64 in the reference's layouts (half spectra
[nflds, N, N, N/2+1] complex; real fields [nflds, N, N, N])."""
import ctypes as C
import os

import numpy as np

from .abi import (PscConfig, PscModel, PscIcParams, PscAdjointSpec, PSC_ENERGY_LEN, PSC_SPECTRA_COLS,
                  PSC_GW_SPECTRA_COLS, PSC_OBJ_CHI2, num_bins)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsc.so")
DEFAULT_MODEL = "preheat"
_libs = {}


def lib_path(model_name=None):
    """libpsc.so holds the default model (preheat); libpsc_<name>.so another
    model header's build."""
    if not model_name or model_name == DEFAULT_MODEL:
        return LIB_PATH
    return os.path.join(_HERE, f"libpsc_{model_name}.so")
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class PscError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"psc error {code}: {msg}")
        self.code = code


def lib(model_name=None):
    path = lib_path(model_name)
    if path not in _libs:
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run __graft_entry__.build() "
                              "or python pspectcosmo_amd/build.py")
        L = C.CDLL(path)
        P = C.POINTER
        vp = C.c_void_p
        L.psc_abi_version.restype = C.c_int
        L.psc_status_string.restype = C.c_char_p
        L.psc_status_string.argtypes = [C.c_int]
        L.psc_num_bins.argtypes = [C.c_int]
        L.psc_generate_ic.argtypes = [P(PscConfig), P(PscModel), P(PscIcParams), _D, _D, _D]
        L.psc_generate_gw_ic.argtypes = [P(PscConfig), P(PscModel), P(PscIcParams), _D, _D]
        L.psc_create.argtypes = [P(PscConfig), P(vp)]
        L.psc_destroy.argtypes = [vp]
        L.psc_last_error.restype = C.c_char_p
        L.psc_last_error.argtypes = [vp]
        L.psc_set_model.argtypes = [vp, C.c_int, P(PscModel)]
        L.psc_load_ic_k.argtypes = [vp, C.c_int, _D, _D, C.c_double, C.c_double]
        L.psc_load_fields.argtypes = [vp, C.c_int, _D, _D, C.c_double, C.c_double, C.c_double]
        L.psc_set_scalars.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.psc_get_scalars.argtypes = [vp, C.c_int, _D]
        L.psc_step.argtypes = [vp, C.c_int]
        L.psc_set_dt.argtypes = [vp, C.c_double]
        L.psc_means.argtypes = [vp, C.c_int, _D]
        L.psc_sync.argtypes = [vp]
        L.psc_get_k.argtypes = [vp, C.c_int, _D, _D]
        L.psc_get_fields.argtypes = [vp, C.c_int, _D, _D]
        L.psc_energies.argtypes = [vp, C.c_int, _D]
        L.psc_energy_densities.argtypes = [vp, C.c_int, _D, _D, _D, _D]
        L.psc_energy_densities_noalloc.argtypes = [vp, C.c_int, _D, _D, _D, _D]
        L.psc_spectra.argtypes = [vp, C.c_int, C.c_int, _D, C.c_int]
        L.psc_load_gw_k.argtypes = [vp, C.c_int, _D, _D]
        L.psc_generate_ic_device.argtypes = [vp, C.c_int, C.POINTER(PscIcParams), _D]
        L.psc_generate_gw_ic_device.argtypes = [vp, C.c_int, C.POINTER(PscIcParams)]
        L.psc_get_gw_k.argtypes = [vp, C.c_int, _D, _D]
        L.psc_gw_spectrum.argtypes = [vp, C.c_int, C.c_int, _D, C.c_int, _D]
        L.psc_gw_omega_log.argtypes = [vp, C.c_int, C.c_int, C.c_int, _D, _D]
        L.psc_set_kernel_timing.argtypes = [vp, C.c_int]
        L.psc_reset_kernel_times.argtypes = [vp]
        L.psc_get_kernel_times.argtypes = [vp, C.c_char_p, C.c_int, _D, P(C.c_longlong), _D,
                                           C.c_int, P(C.c_int)]
        L.psc_device_bytes.argtypes = [vp, _D]
        L.psc_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.c_int]
        L.psc_shape_name.restype = C.c_char_p
        L.psc_shape_name.argtypes = [vp]
        L.psc_list_shapes.argtypes = [vp, C.c_char_p, C.c_int]
        L.psc_set_shape.argtypes = [vp, C.c_char_p]
        L.psc_adjoint.argtypes = [vp, P(PscAdjointSpec), _D, vp]
        L.psc_model_name.restype = C.c_char_p
        L.psc_model_description.restype = C.c_char_p
        L.psc_num_potential_terms.argtypes = [C.c_int]
        L.psc_potential_terms.argtypes = [vp, C.c_int, _D]
        L.psc_set_output_times.argtypes = [vp, C.c_double, C.c_double]
        L.psc_get_step_count.argtypes = [vp, C.c_int, P(C.c_longlong)]
        _libs[path] = L
    return _libs[path]


def exported_symbols():
    """Names declared in include/psc.h (checked against the built library)."""
    return ["psc_abi_version", "psc_status_string", "psc_generate_ic", "psc_generate_gw_ic",
            "psc_create", "psc_destroy", "psc_last_error", "psc_set_model", "psc_load_ic_k",
            "psc_set_scalars", "psc_get_scalars", "psc_step", "psc_sync", "psc_get_k",
            "psc_get_fields", "psc_energies", "psc_spectra", "psc_load_gw_k", "psc_get_gw_k",
            "psc_gw_spectrum", "psc_gw_omega_log", "psc_num_bins", "psc_set_kernel_timing", "psc_get_kernel_times",
            "psc_reset_kernel_times", "psc_device_bytes", "psc_autotune", "psc_shape_name",
            "psc_list_shapes", "psc_set_shape", "psc_adjoint", "psc_set_dt", "psc_means", "psc_load_fields",
            "psc_energy_densities", "psc_generate_ic_device", "psc_generate_gw_ic_device",
            "psc_model_name", "psc_model_description", "psc_num_potential_terms", "psc_potential_terms",
            "psc_set_output_times", "psc_get_step_count", "psc_energy_densities_noalloc"]


def generate_ic(cfg, model, ic, model_name=None):
    """Host IC generator of the library (initialize.cpp:562-810): half spectra
    and (a, ad, add_printed, H)."""
    N, nf = cfg.N, cfg.nflds
    fk = np.zeros((nf, N, N, N // 2 + 1), dtype=np.complex128)
    pk = np.zeros_like(fk)
    sc = np.zeros(4)
    rc = lib(model_name).psc_generate_ic(C.byref(cfg), C.byref(model), C.byref(ic),
                                         fk.view(np.float64).ravel(), pk.view(np.float64).ravel(), sc)
    if rc:
        raise PscError(rc, lib(model_name).psc_status_string(rc).decode())
    return fk, pk, sc


def generate_gw_ic(cfg, model, ic, model_name=None):
    N = cfg.N
    u = np.zeros((6, N, N, N // 2 + 1), dtype=np.complex128)
    pu = np.zeros_like(u)
    rc = lib(model_name).psc_generate_gw_ic(C.byref(cfg), C.byref(model), C.byref(ic),
                                            u.view(np.float64).ravel(), pu.view(np.float64).ravel())
    if rc:
        raise PscError(rc, lib(model_name).psc_status_string(rc).decode())
    return u, pu


class Lattice:
    """A device context holding cfg.batch lattices (psc_ctx)."""

    def __init__(self, cfg, model_name=None):
        self.cfg = cfg
        self.model_name = model_name or DEFAULT_MODEL
        self._L = lib(model_name)
        self.N, self.nf, self.B = cfg.N, cfg.nflds, cfg.batch
        h = C.c_void_p()
        rc = self._L.psc_create(C.byref(cfg), C.byref(h))
        if rc:
            raise PscError(rc, "psc_create: " + self._L.psc_status_string(rc).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.psc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise PscError(rc, self._L.psc_last_error(self._h).decode())

    def set_model(self, member, model):
        self._check(self._L.psc_set_model(self._h, member, C.byref(model)))

    def load_ic_k(self, member, fk, pk, a, ad):
        fk = np.ascontiguousarray(fk, dtype=np.complex128)
        pk = np.ascontiguousarray(pk, dtype=np.complex128)
        self._check(self._L.psc_load_ic_k(self._h, member, fk.view(np.float64).ravel(),
                                        pk.view(np.float64).ravel(), a, ad))

    def generate_ic_device(self, member, ic):
        """Initial conditions generated on the device for one member; returns
        (a, ad, add_printed, H) as generate_ic does."""
        sc = np.zeros(4)
        self._check(self._L.psc_generate_ic_device(self._h, member, C.byref(ic), sc))
        return sc

    def generate_gw_ic_device(self, member, ic):
        self._check(self._L.psc_generate_gw_ic_device(self._h, member, C.byref(ic)))

    def load_fields(self, member, f, pi, t, a, ad):
        f = np.ascontiguousarray(f, dtype=np.float64)
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        self._check(self._L.psc_load_fields(self._h, member, f.ravel(), pi.ravel(), t, a, ad))

    def set_scalars(self, member, t, a, ad, add):
        self._check(self._L.psc_set_scalars(self._h, member, t, a, ad, add))

    def scalars(self, member=0):
        s = np.zeros(4)
        self._check(self._L.psc_get_scalars(self._h, member, s))
        return s

    def step(self, n=1):
        self._check(self._L.psc_step(self._h, int(n)))

    def set_output_times(self, t_out, dt_cap):
        self._check(self._L.psc_set_output_times(self._h, float(t_out), float(dt_cap)))

    def step_count(self, member=0):
        n = C.c_longlong()
        self._check(self._L.psc_get_step_count(self._h, member, C.byref(n)))
        return n.value

    def set_dt(self, dt):
        self._check(self._L.psc_set_dt(self._h, float(dt)))

    def means(self, member=0):
        """Per field: <f>, a^{alpha-3}<pi>, <f^2>, a^{2(alpha-3)}<pi^2>, sigma_f, sigma_fd."""
        out = np.zeros(12)
        self._check(self._L.psc_means(self._h, member, out))
        return out.reshape(2, 6)[: self.nf]

    def sync(self):
        self._check(self._L.psc_sync(self._h))

    def get_k(self, member=0):
        N = self.N
        fk = np.zeros((self.nf, N, N, N // 2 + 1), dtype=np.complex128)
        pk = np.zeros_like(fk)
        self._check(self._L.psc_get_k(self._h, member, fk.view(np.float64).ravel(),
                                    pk.view(np.float64).ravel()))
        return fk, pk

    def fields(self, member=0):
        N = self.N
        f = np.zeros((self.nf, N, N, N))
        p = np.zeros_like(f)
        self._check(self._L.psc_get_fields(self._h, member, f.ravel(), p.ravel()))
        return f, p

    def energy_densities(self, member=0):
        """Snapshot energy densities (pot, kin, grad, tot), each (N, N, N)."""
        N = self.N
        out = [np.zeros((N, N, N)) for _ in range(4)]
        self._check(self._L.psc_energy_densities(self._h, member, *[o.ravel() for o in out]))
        return tuple(out)

    def energy_densities_noalloc(self, member=0):
        """The same densities through the library's fixed staging array (no
        device allocation in the call)."""
        N = self.N
        out = [np.zeros((N, N, N)) for _ in range(4)]
        self._check(self._L.psc_energy_densities_noalloc(self._h, member, *[o.ravel() for o in out]))
        return tuple(out)

    def energies(self, member=0):
        e = np.zeros(PSC_ENERGY_LEN)
        self._check(self._L.psc_energies(self._h, member, e))
        return e

    def spectra(self, fld, member=0):
        nb = num_bins(self.N)
        out = np.zeros((nb, PSC_SPECTRA_COLS))
        self._check(self._L.psc_spectra(self._h, member, fld, out.ravel(), nb))
        return out

    def load_gw_k(self, member, u, pu):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        pu = np.ascontiguousarray(pu, dtype=np.complex128)
        self._check(self._L.psc_load_gw_k(self._h, member, u.view(np.float64).ravel(),
                                        pu.view(np.float64).ravel()))

    def get_gw_k(self, member=0):
        N = self.N
        u = np.zeros((6, N, N, N // 2 + 1), dtype=np.complex128)
        pu = np.zeros_like(u)
        self._check(self._L.psc_get_gw_k(self._h, member, u.view(np.float64).ravel(),
                                       pu.view(np.float64).ravel()))
        return u, pu

    def gw_spectrum(self, projection, member=0):
        nb = num_bins(self.N)
        out = np.zeros((nb, PSC_GW_SPECTRA_COLS))
        tot = np.zeros(2)
        self._check(self._L.psc_gw_spectrum(self._h, member, projection, out.ravel(), nb, tot))
        return out, tot

    def gw_omega_log(self, projection, nlog=256, member=0):
        kc = np.zeros(nlog)
        om = np.zeros(nlog)
        self._check(self._L.psc_gw_omega_log(self._h, member, projection, nlog, kc, om))
        return kc, om

    def set_kernel_timing(self, on):
        self._check(self._L.psc_set_kernel_timing(self._h, 1 if on else 0))

    def reset_kernel_times(self):
        self._check(self._L.psc_reset_kernel_times(self._h))

    def kernel_times(self):
        cap = 32
        names = C.create_string_buffer(2048)
        ms = np.zeros(cap)
        bytes_ = np.zeros(cap)
        counts = (C.c_longlong * cap)()
        n = C.c_int()
        self._check(self._L.psc_get_kernel_times(self._h, names, 2048, ms, counts, bytes_, cap,
                                               C.byref(n)))
        out = {}
        for i, name in enumerate(names.value.decode().split(",")[: n.value]):
            out[name] = {"ms": float(ms[i]), "count": int(counts[i]), "bytes": float(bytes_[i])}
        return out

    def autotune(self, nsteps=3):
        buf = C.create_string_buffer(4096)
        self._check(self._L.psc_autotune(self._h, int(nsteps), buf, 4096))
        out = {}
        for item in buf.value.decode().split(";"):
            if "=" in item:
                k, v = item.split("=")
                out[k] = float(v)
        return out

    def shape_name(self):
        return self._L.psc_shape_name(self._h).decode()

    def shapes(self):
        buf = C.create_string_buffer(4096)
        self._check(self._L.psc_list_shapes(self._h, buf, 4096))
        return buf.value.decode().split(",")

    def set_shape(self, name):
        self._check(self._L.psc_set_shape(self._h, name.encode()))

    def adjoint(self, nsteps, checkpoint_every, phi0, dphi0, chi0=0.0, want_field=True, objective=0,
                projection=1, coupling=0):
        """Gradient of the final <chi^2> (objective 0) or of the final GW energy
        fraction (objective 1, with `projection`) by psc_adjoint: returns a dict
        with J, dJ/d(coupling) (key "dg2"; g2 by default), dJ/dphi0, dJ/dchi(x) (N^3 array or None), forward and
        backward milliseconds and the device bytes of the adjoint workspace."""
        spec = PscAdjointSpec()
        spec.objective, spec.nsteps, spec.checkpoint_every, spec.member = int(objective), int(nsteps), \
            int(checkpoint_every), 0
        spec.projection = int(projection)
        spec.coupling = int(coupling)
        spec.phi0, spec.dphi0, spec.chi0 = phi0, dphi0, chi0
        out = np.zeros(6)
        dchi = np.zeros((self.N,) * 3) if want_field else None
        ptr = dchi.ctypes.data_as(C.c_void_p) if want_field else None
        self._check(self._L.psc_adjoint(self._h, C.byref(spec), out, ptr))
        return {"J": out[0], "dg2": out[1], "dphi0": out[2], "dchi": dchi, "forward_ms": out[3],
                "backward_ms": out[4], "extra_bytes": out[5]}

    def potential_terms(self, member=0):
        """Lattice averages of the model's potential terms."""
        out = np.zeros(4)
        self._check(self._L.psc_potential_terms(self._h, member, out))
        return out[:self._L.psc_num_potential_terms(self.nf)]

    def device_bytes(self):
        b = np.zeros(1)
        self._check(self._L.psc_device_bytes(self._h, b))
        return float(b[0])
