"""ctypes mirror of include/psc.h (structs, constants) shared by the Python
bindings of the HIP library and by the tests."""
import ctypes as C

PSC_OK = 0
PSC_ERR_ARG = 1
PSC_ERR_DEVICE = 2
PSC_ERR_NOMEM = 3
PSC_ERR_STATE = 4
PSC_ERR_NONFINITE = 5
PSC_ERR_NEGATIVE_A = 6
PSC_ERR_UNSUPPORTED = 7

PSC_VV2 = 2
PSC_Y4 = 4
PSC_FP64 = 0
PSC_FP32 = 1
PSC_EXPANSION_STATIC = 0
PSC_EXPANSION_SELF = 2
PSC_DT_FIXED = 0
PSC_DT_REFERENCE = 1
PSC_GW_PROJ_REFERENCE = 0
PSC_GW_PROJ_TT = 1
PSC_ENERGY_LEN = 16
PSC_SPECTRA_COLS = 5
PSC_GW_SPECTRA_COLS = 5
PSC_OBJ_CHI2 = 0
PSC_OBJ_GW_FRACTION = 1
PSC_COUPLING_G2 = 0
PSC_COUPLING_M2 = 1
PSC_COUPLING_LAMBDA = 2
PSC_COUPLING_C0 = 3
PSC_MAX_TERMS = 4


class PscConfig(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("nflds", C.c_int), ("integrator", C.c_int), ("precision", C.c_int),
        ("batch", C.c_int), ("with_gw", C.c_int), ("alpha", C.c_int), ("expansion", C.c_int),
        ("yoshida_exact", C.c_int), ("use_graph", C.c_int), ("autotune", C.c_int),
        ("pad", C.c_int), ("dt_rule", C.c_int), ("n_pad", C.c_int), ("reserved_i", C.c_int * 2),
        ("kIR", C.c_double), ("dt", C.c_double), ("fStar", C.c_double),
        ("omegaStar", C.c_double), ("MPl", C.c_double), ("reserved_d", C.c_double * 3),
    ]


class PscModel(C.Structure):
    _fields_ = [("m2", C.c_double), ("g2", C.c_double), ("lambda_", C.c_double), ("c", C.c_double * 5)]


class PscIcParams(C.Structure):
    _fields_ = [
        ("f0", C.c_double * 2), ("fd0", C.c_double * 2), ("a0", C.c_double),
        ("seed", C.c_int), ("reserved_i", C.c_int), ("kcutoff", C.c_double), ("reserved_d", C.c_double * 3),
    ]


class PscAdjointSpec(C.Structure):
    _fields_ = [
        ("objective", C.c_int), ("nsteps", C.c_int), ("checkpoint_every", C.c_int), ("member", C.c_int),
        ("projection", C.c_int), ("coupling", C.c_int),
        ("phi0", C.c_double), ("dphi0", C.c_double), ("chi0", C.c_double), ("reserved", C.c_double * 4),
    ]


def make_config(N, nflds=2, integrator=PSC_VV2, precision=PSC_FP64, batch=1, with_gw=0,
                alpha=0, expansion=PSC_EXPANSION_SELF, yoshida_exact=0, use_graph=0, autotune=0,
                kIR=25.0, dt=0.005, fStar=1.0, omegaStar=1.0, MPl=1.0, pad=0, dt_rule=0, n_pad=0):
    c = PscConfig()
    c.N, c.nflds, c.integrator, c.precision = N, nflds, integrator, precision
    c.batch, c.with_gw, c.alpha, c.expansion = batch, with_gw, alpha, expansion
    c.yoshida_exact, c.use_graph, c.autotune = yoshida_exact, use_graph, autotune
    c.kIR, c.dt, c.fStar, c.omegaStar, c.MPl = kIR, dt, fStar, omegaStar, MPl
    c.pad = pad
    c.n_pad = n_pad
    c.dt_rule = dt_rule
    return c


def make_model(m2=1.0, g2=0.0, lam=0.0, c=()):
    """Couplings of psc_model; which of them a model reads is stated by its
    header (preheat: m2, g2; lambda_phi4: lam, g2)."""
    m = PscModel()
    m.m2, m.g2, m.lambda_ = m2, g2, lam
    for i, v in enumerate(c):
        m.c[i] = v
    return m


def make_ic(f0=(1.0, 0.0), fd0=(0.0, 0.0), a0=1.0, seed=1, kcutoff=0.0):
    p = PscIcParams()
    p.f0[0], p.f0[1] = f0
    p.fd0[0], p.fd0[1] = fd0
    p.a0, p.seed, p.kcutoff = a0, seed, kcutoff
    return p


def num_bins(N):
    """maxnumbins - 1 (parameters.h:25): bins 1..maxnumbins-1 are reported."""
    import math
    return int(math.sqrt(3.0) * (N // 2)) + 2 - 1


def half_len(N):
    return N * N * (N // 2 + 1)
