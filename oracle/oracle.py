"""TEST INFRASTRUCTURE ONLY: numpy/ctypes front end of the CPU oracle (oracle.cc).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the
checker. Arrays are (8, mz, my, mx) numpy arrays in the reference's x-fastest layout
(ref acc-runtime/api/acc_runtime.h:58-59), fields in the order lnrho, ux, uy, uz, ax, ay, az, ss.

Also restated here (pure numpy):
  * acGetError / acEvalErrorWithMaximumError / acVerifyMesh, the ulp/abs/rel error metric of
    ref src/utils/verification.cc:44-129,199-269.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NF = 8


class OracleConfig(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("ds", ctypes.c_double * 3),
                ("cs_sound", ctypes.c_double), ("nu_visc", ctypes.c_double),
                ("zeta", ctypes.c_double), ("eta", ctypes.c_double), ("mu0", ctypes.c_double),
                ("cp_sound", ctypes.c_double), ("gamma", ctypes.c_double),
                ("lnT0", ctypes.c_double), ("lnrho0", ctypes.c_double),
                ("K_heatcond", ctypes.c_double),
                ("lupwd", ctypes.c_int), ("lentropy", ctypes.c_int)]


# Defaults of ref config/astaroth.conf (physics) for the parameters the oracle uses.
DEFAULTS = dict(cs_sound=1.0, nu_visc=2e-5, zeta=0.01, eta=5e-5, mu0=1.4, cp_sound=1.0,
                gamma=0.5, lnT0=1.2, lnrho0=1.3, K_heatcond=1e-3, lupwd=1, lentropy=1)


def make_config(n, ds, **params):
    p = dict(DEFAULTS)
    p.update(params)
    c = OracleConfig()
    c.nx, c.ny, c.nz = n
    if np.isscalar(ds):
        ds = (ds, ds, ds)
    for i in range(3):
        c.ds[i] = float(ds[i])
    for k in DEFAULTS:
        setattr(c, k, p[k])
    return c


def config_from_info(info):
    """OracleConfig from an astaroth_amd AcMeshInfo (by parameter name)."""
    return make_config(info["AC_nlocal"], info["AC_ds"],
                       cs_sound=float(info["AC_cs_sound"]), nu_visc=float(info["AC_nu_visc"]),
                       zeta=float(info["AC_zeta"]), eta=float(info["AC_eta"]),
                       mu0=float(info["AC_mu0"]), cp_sound=float(info["AC_cp_sound"]),
                       gamma=float(info["AC_gamma"]), lnT0=float(info["AC_lnT0"]),
                       lnrho0=float(info["AC_lnrho0"]), K_heatcond=float(info["AC_K_heatcond"]),
                       lupwd=int(info["AC_lupwd"]), lentropy=int(info["AC_lentropy"]))


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        pp_d = P(P(ctypes.c_double))
        pp_f = P(P(ctypes.c_float))
        cfg = P(OracleConfig)
        L.oracle_integrate_step_f64.argtypes = [pp_d, cfg, ctypes.c_double]
        L.oracle_integrate_step_f32.argtypes = [pp_f, cfg, ctypes.c_double]
        L.oracle_periodic_fill_f64.argtypes = [pp_d, ctypes.c_int, cfg]
        L.oracle_periodic_fill_f32.argtypes = [pp_f, ctypes.c_int, cfg]
        L.oracle_rhs_f64.argtypes = [pp_d, cfg, pp_d]
        L.oracle_reduce_f64.argtypes = [pp_d, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, cfg]
        L.oracle_reduce_f64.restype = ctypes.c_longdouble
        L.oracle_reduce_f32.argtypes = [pp_f, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, cfg]
        L.oracle_reduce_f32.restype = ctypes.c_longdouble
        L.refcpu_integrate_step_f64.argtypes = [pp_d, pp_d, cfg, ctypes.c_double, ctypes.c_int]
        L.refcpu_integrate_step_f32.argtypes = [pp_f, pp_f, cfg, ctypes.c_double, ctypes.c_int]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_seed_rng.argtypes = [ctypes.c_uint32]
        L.oracle_seed_rng.restype = None
        L.oracle_mt19937_draw.restype = ctypes.c_uint32
        L.oracle_shell_size.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oracle_helical_forcing_params.argtypes = [ctypes.c_double] * 3 + [P(ctypes.c_double)]
        L.oracle_add_forcing_f64.argtypes = [pp_d] + [ctypes.c_int] * 3 + [P(ctypes.c_double)]
        L.oracle_add_forcing_f32.argtypes = [pp_f] + [ctypes.c_int] * 3 + [P(ctypes.c_double)]
        _lib = L
    return _lib


def _ptrs(arr):
    assert arr.flags["C_CONTIGUOUS"]
    ct = ctypes.c_double if arr.dtype == np.float64 else ctypes.c_float
    P = ctypes.POINTER(ct)
    n = arr.shape[0]
    a = (P * n)(*[arr[q].ctypes.data_as(P) for q in range(n)])
    return a


def integrate_step(fields, cfg, dt):
    """acHostIntegrateStep (ref modelsolver.cc:1214-1272) in place on `fields` (float32/64)."""
    if fields.dtype == np.float64:
        lib().oracle_integrate_step_f64(_ptrs(fields), ctypes.byref(cfg), dt)
    elif fields.dtype == np.float32:
        lib().oracle_integrate_step_f32(_ptrs(fields), ctypes.byref(cfg), dt)
    else:
        raise TypeError(fields.dtype)


def periodic_fill(fields, cfg):
    """acHostMeshApplyPeriodicBounds (ref memory.cc:75-149) in place."""
    fn = lib().oracle_periodic_fill_f64 if fields.dtype == np.float64 else lib().oracle_periodic_fill_f32
    fn(_ptrs(fields), fields.shape[0], ctypes.byref(cfg))


def rhs(fields, cfg):
    """Right-hand side of all 8 fields at interior cells (long double, rounded to double).
    Halos of `fields` (float64) must be filled."""
    out = np.zeros_like(fields, dtype=np.float64)
    lib().oracle_rhs_f64(_ptrs(np.ascontiguousarray(fields, dtype=np.float64)), ctypes.byref(cfg),
                         _ptrs(out))
    return out


def reduce(bufs, cfg, map_kind, op, rms):
    """Long double host reduction (ref modelreduce.cc:108-340) over the interior.
    bufs: list of 1-4 (mz, my, mx) arrays."""
    arr = np.ascontiguousarray(np.stack(bufs))
    fn = lib().oracle_reduce_f64 if arr.dtype == np.float64 else lib().oracle_reduce_f32
    return fn(_ptrs(arr), arr.shape[0], map_kind, op, int(rms), ctypes.byref(cfg))


def refcpu_step(fields, wbuf, cfg, dt, nthreads=0):
    """One RK3 step of the ref-cpu algorithm (two-pass, OpenMP) in the arithmetic of `fields`:
    float64 is the reference CPU build's default, float32 the same build with AcReal = float."""
    assert fields.dtype == wbuf.dtype
    if fields.dtype == np.float64:
        lib().refcpu_integrate_step_f64(_ptrs(fields), _ptrs(wbuf), ctypes.byref(cfg), dt, nthreads)
    elif fields.dtype == np.float32:
        lib().refcpu_integrate_step_f32(_ptrs(fields), _ptrs(wbuf), ctypes.byref(cfg), dt, nthreads)
    else:
        raise TypeError(fields.dtype)


def num_threads():
    return lib().oracle_num_threads()


# ---------------------------------------------------------------- helical forcing
def seed_rng(seed):
    """seed_rng of ref src/utils/astaroth_random.cc on the oracle's own std::mt19937."""
    lib().oracle_seed_rng(seed)


def mt19937_draw():
    return int(lib().oracle_mt19937_draw())


def shell_size(kmin, kmax):
    return int(lib().oracle_shell_size(kmin, kmax))


def helical_forcing_params(relhel, kmin, kmax):
    """generateHelicalForcingParams (ref src/utils/helical_forcing.cc:325-341) in FP64:
    dict(k, f_re, f_im (3-vectors), phase, kaver), before rounding to AcReal."""
    out = (ctypes.c_double * 11)()
    if lib().oracle_helical_forcing_params(relhel, kmin, kmax, out) != 0:
        raise ValueError("empty wave-vector shell")
    v = list(out)
    return dict(k=np.array(v[0:3]), f_re=np.array(v[3:6]), f_im=np.array(v[6:9]), phase=v[9],
                kaver=v[10])


def add_forcing(fields, cfg, k, f_re, f_im, phase, magnitude, kaver, cs2_sound, ngrid, dt,
                current_time, switch_forcing):
    """The substep-2 forcing of ref mhdsolver.ac:286-331,614-631 added to fields[1:4] in place
    (long double, rounded to the fields' dtype). `cfg` supplies the interior size and ds."""
    p = (ctypes.c_double * 22)(*(list(map(float, k)) + list(map(float, f_re)) +
                                 list(map(float, f_im)) +
                                 [float(phase), float(magnitude), float(kaver), float(cs2_sound)] +
                                 [cfg.ds[0], cfg.ds[1], cfg.ds[2]] + list(map(float, ngrid)) +
                                 [float(dt), float(current_time), float(switch_forcing)]))
    uu = np.ascontiguousarray(fields[1:4])
    if fields.dtype == np.float64:
        lib().oracle_add_forcing_f64(_ptrs(uu), cfg.nx, cfg.ny, cfg.nz, p)
    else:
        lib().oracle_add_forcing_f32(_ptrs(uu), cfg.nx, cfg.ny, cfg.nz, p)
    fields[1:4] = uu


# ---------------------------------------------------------------- verification metric
def get_error(model, candidate, real_bits):
    """acGetError (ref verification.cc:44-81) for scalars; real_bits = 24 (float) or 53."""
    model = np.longdouble(model)
    candidate = np.longdouble(candidate)
    if model == candidate:
        return dict(abs=0.0, ulp=0.0, rel=0.0)
    if not (np.isfinite(model) and np.isfinite(candidate)):
        return dict(abs=np.inf, ulp=np.inf, rel=np.inf)
    e = np.floor(np.log(np.abs(model)) / np.log(np.longdouble(2)))
    ulp = np.power(np.longdouble(2), e - (real_bits - 1))
    eps = np.longdouble(0.5) * np.power(np.longdouble(2), -(real_bits - 1))
    abs_err = np.abs(model - candidate)
    return dict(abs=float(abs_err), ulp=float(abs_err / ulp),
                rel=float(np.abs(1 - candidate / model) / eps))


def verify_field(model, candidate, interior, max_ulp, real_bits):
    """Per-field check of ref verification.cc:94-129,199-269: the error at the point of the
    maximum absolute error, accepted if ulp < max_ulp or abs < min|model| * AC_REAL_EPSILON."""
    m = model[interior].astype(np.longdouble)
    c = candidate[interior].astype(np.longdouble)
    diff = np.abs(m - c)
    idx = np.unravel_index(int(np.argmax(diff)), diff.shape)
    err = get_error(m[idx], c[idx], real_bits)
    eps = np.finfo(np.float32 if real_bits == 24 else np.float64).eps
    min_mag = float(np.min(np.abs(m)))
    ok = err["ulp"] < max_ulp or err["abs"] < min_mag * eps
    err.update(ok=bool(ok), where=tuple(int(v) for v in idx), min_mag=min_mag,
               max_mag=float(np.max(np.abs(m))))
    return err


def interior_slices(shape_zyx, nghost=3):
    mz, my, mx = shape_zyx
    return (slice(nghost, mz - nghost), slice(nghost, my - nghost), slice(nghost, mx - nghost))


def field_errors(model, candidate, nghost=3):
    """Per-field max-relative (normalised by the field's max |model|) and relative L2 errors."""
    out = []
    sl = interior_slices(model.shape[1:], nghost)
    for q in range(model.shape[0]):
        m = model[q][sl].astype(np.float64)
        c = candidate[q][sl].astype(np.float64)
        scale = max(float(np.max(np.abs(m))), 1e-300)
        maxrel = float(np.max(np.abs(m - c))) / scale
        l2 = float(np.sqrt(np.sum((m - c) ** 2)) / max(np.sqrt(np.sum(m ** 2)), 1e-300))
        out.append(dict(maxrel=maxrel, l2=l2, scale=scale))
    return out
