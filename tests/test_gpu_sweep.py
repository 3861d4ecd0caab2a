"""GPU parity of the unrolled sweep kernel (solve_sweep.hip, FP32, 64 x 4 column blocks): against
the oracle at small sizes for several z-chunks and every physics switch, against the delay-line
kernel at 128^3, and the periodic-image multi-step path against fill + substep + swap bit for bit
for every compensation width. The kernel is pinned with AC_solve_tile_x / AC_solve_tile_y /
AC_solve_zchunk and AC_solve_kernel = 1."""
import ctypes
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac
import oracle.oracle as O

pytestmark = pytest.mark.gpu

SWEEP = 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def physics_info(lib, n, shape=(0, 0, 0), kind=0, **over):
    p = dict(AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0, AC_nu_visc=5e-3,
             AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3, AC_lupwd=True, AC_lentropy=True,
             AC_solve_tile_x=shape[0], AC_solve_tile_y=shape[1], AC_solve_zchunk=shape[2],
             AC_solve_kernel=kind)
    p.update(over)
    return lib.config(n, **p)


def initial(lib, info, seed=2024):
    mesh = ac.HostMesh(lib, info)
    assert lib.acHostMeshInitFourier(ctypes.byref(mesh.struct), seed, 16, 4, 0.05, 0.0) == 0
    data = mesh.data.copy()
    rng = np.random.default_rng(seed)
    data[0] += 0.01 * rng.standard_normal(data[0].shape).astype(data.dtype)
    data[7] += 0.01 * rng.standard_normal(data[7].shape).astype(data.dtype)
    O.periodic_fill(data, O.config_from_info(info))
    return data


def run(lib, info, data, dt, steps, expect=None):
    mesh = ac.HostMesh(lib, info, data)
    with ac.Device(lib, info) as dev:
        if expect is not None:
            sh = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int())
            lib.acDeviceGetSolveLaunchShape(dev.handle, *[ctypes.byref(v) for v in sh])
            kind = ctypes.c_int()
            lib.acDeviceGetSolveKernel(dev.handle, ctypes.byref(kind))
            assert (tuple(v.value for v in sh), kind.value) == expect
        dev.load_mesh(mesh)
        for _ in range(steps):
            dev.step(dt)
        dev.periodic()
        out = ac.HostMesh(lib, info)
        dev.store_mesh(out)
    return out.data


@pytest.fixture(scope="module")
def lib(require_gpu):
    return ac.load("fp32")


@pytest.mark.parametrize("n,zc", [((64, 32, 32), 32), ((64, 16, 64), 16), ((128, 8, 24), 8),
                                  ((64, 4, 12), 12)])
@pytest.mark.parametrize("upwd,entropy", [(True, True), (False, False), (True, False), (False, True)])
def test_sweep_vs_oracle(lib, n, zc, upwd, entropy):
    shape = (64, 4, zc)
    say(f"case sweep vs oracle n={n} shape={shape} upwd={upwd} entropy={entropy}")
    info = physics_info(lib, n, shape, SWEEP, AC_lupwd=upwd, AC_lentropy=entropy)
    data = initial(lib, info)
    dt, steps = 2e-3, 3
    cand = run(lib, info, data, dt, steps, expect=(shape, SWEEP))
    model = data.astype(np.float64)
    cfg = O.config_from_info(info)
    for _ in range(steps):
        O.integrate_step(model, cfg, dt)
    O.periodic_fill(model, cfg)
    for q, e in enumerate(O.field_errors(model, cand)):
        say(f"  {ac.FIELD_NAMES[q]}: maxrel={e['maxrel']:.3g}")
        assert e["maxrel"] < 2e-5, (q, e)


@pytest.mark.parametrize("zc", [32, 64, 128])
def test_sweep_vs_delay_line_128(lib, zc):
    """Same arithmetic up to the order of a few sums: the two kernels agree to ~1e-6 of the field
    scale at 128^3; an index or coefficient slip is >= 1e-3."""
    say(f"case sweep vs delay-line 128^3 zc={zc}")
    n = (128, 128, 128)
    info_s = physics_info(lib, n, (64, 4, zc), SWEEP)
    info_t = physics_info(lib, n, (64, 4, zc), 0)
    data = initial(lib, info_s, seed=5)
    dt, steps = 1e-2, 2
    a = run(lib, info_s, data, dt, steps, expect=((64, 4, zc), SWEEP))
    b = run(lib, info_t, data, dt, steps, expect=((64, 4, zc), 0))
    for q, e in enumerate(O.field_errors(b, a)):
        say(f"  {ac.FIELD_NAMES[q]}: maxrel={e['maxrel']:.3g}")
        assert e["maxrel"] < 1e-5, (q, e)


def test_sweep_anisotropic_falls_back(lib):
    """On a grid with unequal spacing the sweep (which shares one coefficient set between the
    axes) does not run; the pinned shape falls back to the direct kernel and the result still
    matches the oracle."""
    n = (64, 16, 32)
    info = physics_info(lib, n, (64, 4, 16), SWEEP, AC_len=(2 * np.pi, np.pi, 2 * np.pi))
    data = initial(lib, info, seed=3)
    dt, steps = 1e-3, 2
    cand = run(lib, info, data, dt, steps)
    model = data.astype(np.float64)
    cfg = O.config_from_info(info)
    for _ in range(steps):
        O.integrate_step(model, cfg, dt)
    O.periodic_fill(model, cfg)
    for q, e in enumerate(O.field_errors(model, cand)):
        assert e["maxrel"] < 2e-5, (q, e)


@pytest.mark.parametrize("kb", [0, 8, 16, 32])
@pytest.mark.parametrize("mode", ["courant", "reference"])
@pytest.mark.parametrize("n,zc", [((64, 16, 32), 16), ((128, 8, 16), 8)])
def test_sweep_fused_periodic_steps(lib, n, zc, kb, mode):
    """acDeviceIntegrateStepsCFL with the sweep reads the periodic images of the halo cells (no
    fill before a substep); the result equals fill + substep + swap, bit for bit, and so does
    the next timestep."""
    shape = (64, 4, zc)
    say(f"case sweep fused periodic steps n={n} shape={shape} kb={kb} mode={mode}")
    info = physics_info(lib, n, shape, SWEEP, AC_kahan_bits=kb)
    data = initial(lib, info, seed=11)
    courant = lib.np_real(0.3)
    tmode = ac.AC_TIMESTEP_COURANT if mode == "courant" else ac.AC_TIMESTEP_REFERENCE
    outs = []
    for fused in (True, False):
        mesh = ac.HostMesh(lib, info, data)
        with ac.Device(lib, info) as dev:
            dev.load_mesh(mesh)
            ac.check(lib.acDeviceSetTimestepMode(dev.handle, tmode), "mode")
            dev.periodic()
            ac.check(lib.acDeviceUpdateTimestepCFL(dev.handle, 0, courant), "cfl")
            if fused:
                ac.check(lib.acDeviceIntegrateStepsCFL(dev.handle, 0, 2, courant, None), "steps")
            else:
                for _ in range(2):
                    for s in range(3):
                        dev.periodic()
                        ac.check(lib.acDeviceIntegrateSubstepDeviceDt(dev.handle, 0, s, courant),
                                 "substep")
                        dev.swap()
            dev.periodic()
            out = ac.HostMesh(lib, info)
            dev.store_mesh(out)
            dt = ctypes.c_double()
            ac.check(lib.acDeviceGetTimestep(dev.handle, 0, ctypes.byref(dt)), "dt")
            outs.append((out.data.copy(), dt.value))
    diff = np.argwhere(outs[0][0] != outs[1][0])
    say(f"  differing values: {len(diff)}  dt {outs[0][1]:.17g} {outs[1][1]:.17g}")
    assert len(diff) == 0, diff[:5]
    assert outs[0][1] == outs[1][1]
    assert np.isfinite(outs[0][0]).all()
    if mode == "courant":
        # the next dt follows max |u| of the new state (reduced in FP64 by the substep-2 kernel)
        sl = O.interior_slices(outs[0][0].shape[1:])
        u2 = sum(outs[0][0][q][sl].astype(np.float64) ** 2 for q in (1, 2, 3))
        want = float(courant) * float(info["AC_dsmin"]) / (1.0 + np.sqrt(u2.max()))
        assert outs[0][1] == pytest.approx(want, rel=1e-12)
