"""The Kahan-compensated FP32 kernels against the long-double oracle (FP64 storage), on the plain
method's CPU-runnable case: 16 seeded Fourier modes with |k| <= 4 and amplitude 0.05, dt = 1e-3,
10 steps, upwinding and entropy on. 32^3 runs the direct kernel (the case itself); 64 x 32 x 32
runs the 64 x 4 and 64 x 2 delay-line kernels and the sweep with the same recipe. Each
compensation width (8, 16, 32 bits) is held to a per-field bound, and the compensated error of the
fields that start at zero and stay small (lnrho, ss: their error is the rounding of the state
update, which the compensation carries) must be below plain FP32's."""
import ctypes
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac
import oracle.oracle as O

pytestmark = pytest.mark.gpu

SHAPES = {"direct": ((32, 32, 32), (-1, 0, 0), 0), "64x4": ((64, 32, 32), (64, 4, 32), 0),
          "64x2": ((64, 32, 32), (64, 2, 32), 0), "sweep": ((64, 32, 32), (64, 4, 32), 1)}

# per-field max-relative bounds against the oracle (normalised by the field's largest |value|):
# lnrho, u (3), A (3), ss
BOUND = {0: 2e-5, 8: 1e-5, 16: 1e-5, 32: 5e-6}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def c1_info(lib, n, shape, kind, kb):
    return lib.config(n, AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0,
                      AC_nu_visc=5e-3, AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3,
                      AC_lupwd=True, AC_lentropy=True, AC_kahan_bits=kb,
                      AC_solve_tile_x=shape[0], AC_solve_tile_y=shape[1],
                      AC_solve_zchunk=shape[2], AC_solve_kernel=kind)


def run(lib, info, data, dt, steps):
    mesh = ac.HostMesh(lib, info, data)
    with ac.Device(lib, info) as dev:
        dev.load_mesh(mesh)
        for _ in range(steps):
            dev.step(dt)
        dev.periodic()
        out = ac.HostMesh(lib, info)
        dev.store_mesh(out)
    return out.data


@pytest.fixture(scope="module")
def oracle_cache():
    return {}


@pytest.mark.parametrize("where", list(SHAPES))
def test_kahan_vs_oracle_c1(where, oracle_cache, require_gpu):
    lib = ac.load("fp32")
    n, shape, kind = SHAPES[where]
    dt, steps = np.float32(1e-3), 10
    info0 = c1_info(lib, n, shape, kind, 0)
    mesh = ac.HostMesh(lib, info0)
    assert lib.acHostMeshInitFourier(ctypes.byref(mesh.struct), 2024, 16, 4, 0.05, 0.0) == 0
    data = mesh.data.copy()
    if n not in oracle_cache:
        model = data.astype(np.float64)
        cfg = O.config_from_info(info0)
        for _ in range(steps):
            O.integrate_step(model, cfg, float(dt))
        O.periodic_fill(model, cfg)
        oracle_cache[n] = model
    model = oracle_cache[n]
    errs = {}
    for kb in (0, 8, 16, 32):
        cand = run(lib, c1_info(lib, n, shape, kind, kb), data, dt, steps)
        errs[kb] = O.field_errors(model, cand)
        say(f"case kahan vs oracle {where} n={n} kb={kb}: " +
            " ".join(f"{ac.FIELD_NAMES[q][7:]}={e['maxrel']:.2e}" for q, e in enumerate(errs[kb])))
    for kb in (0, 8, 16, 32):
        for q, e in enumerate(errs[kb]):
            assert e["maxrel"] < BOUND[kb], (where, kb, q, e)
            assert e["l2"] < BOUND[kb], (where, kb, q, e)
    for kb in (8, 16, 32):
        for q in (0, 7):
            assert errs[kb][q]["maxrel"] < errs[0][q]["maxrel"], (where, kb, q)
