"""Helical forcing on the device (ref mhdsolver.ac:286-331, added to u after the final update of
substep 2, :614-631): one RK3 step with forcing against the oracle, the forcing term's Fourier
transform on a uniform state (where the right-hand side vanishes, so u after one step is the
force itself), the AC_switch_forcing threshold, and the multi-step device-dt path against the
step-by-step one."""
import ctypes
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac
import oracle.oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=["fp64", "fp32"])
def lib(request, require_gpu):
    return ac.load(request.param)


def forced_info(lib, n, shape=(0, 0, 0), relhel=0.5, magnitude=0.05, seed=11, **over):
    p = dict(AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0, AC_nu_visc=5e-3,
             AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3, AC_lupwd=True, AC_lentropy=True,
             AC_lforcing=True, AC_switch_forcing=-1.0, AC_current_time=0.0,
             AC_solve_tile_x=shape[0], AC_solve_tile_y=shape[1], AC_solve_zchunk=shape[2])
    p.update(over)
    info = lib.config(n, **p)
    lib.seed_rng(seed)
    fp = lib.generateHelicalForcingParams(lib.np_real(relhel), lib.np_real(magnitude),
                                          lib.np_real(1.0), lib.np_real(3.0))
    assert lib.loadForcingParamsToMeshInfo(fp, ctypes.byref(info)) == 0
    return info, fp


def initial(lib, info, seed=5):
    mesh = ac.HostMesh(lib, info)
    assert lib.acHostMeshInitFourier(ctypes.byref(mesh.struct), seed, 16, 4, 0.05, 1e-3) == 0
    return mesh.data.copy()


def oracle_step_with_forcing(lib, info, data, fp, dt):
    cfg = O.config_from_info(info)
    model = data.astype(np.float64)
    O.integrate_step(model, cfg, dt)
    ng = info["AC_ngrid"]
    O.add_forcing(model, cfg, (fp.k_force.x, fp.k_force.y, fp.k_force.z),
                  (fp.ff_hel_re.x, fp.ff_hel_re.y, fp.ff_hel_re.z),
                  (fp.ff_hel_im.x, fp.ff_hel_im.y, fp.ff_hel_im.z), fp.phase, fp.magnitude,
                  fp.kaver, float(info["AC_cs2_sound"]), tuple(ng), dt,
                  float(info["AC_current_time"]), float(info["AC_switch_forcing"]))
    O.periodic_fill(model, cfg)
    return model


def run_steps(lib, info, data, dt, steps):
    mesh = ac.HostMesh(lib, info, data)
    with ac.Device(lib, info) as dev:
        dev.load_mesh(mesh)
        for _ in range(steps):
            dev.step(lib.np_real(dt))
        dev.periodic()
        out = ac.HostMesh(lib, info)
        dev.store_mesh(out)
    return out.data.copy()


@pytest.mark.parametrize("n,shape", [((16, 16, 16), (-1, 0, 0)), ((64, 8, 16), (64, 4, 16)),
                                     ((32, 32, 32), (32, 8, 32))])
@pytest.mark.parametrize("relhel", [0.0, 1.0])
def test_forced_step_vs_oracle(lib, n, shape, relhel):
    say(f"case forced step vs oracle {lib.precision} n={n} shape={shape} relhel={relhel}")
    info, fp = forced_info(lib, n, shape, relhel=relhel)
    data = initial(lib, info)
    dt = 2e-3
    cand = run_steps(lib, info, data, dt, 1)
    model = oracle_step_with_forcing(lib, info, data, fp, dt)
    tol = 2e-5 if lib.precision == "fp32" else 1e-12
    for q, e in enumerate(O.field_errors(model, cand)):
        say(f"  {ac.FIELD_NAMES[q]}: maxrel={e['maxrel']:.3g}")
        assert e["maxrel"] < tol, (q, e)
    # the forcing moved u by far more than the tolerance
    plain = data.astype(np.float64)
    O.integrate_step(plain, O.config_from_info(info), dt)
    O.periodic_fill(plain, O.config_from_info(info))
    assert max(np.abs(model[q] - plain[q]).max() for q in (1, 2, 3)) > 100 * tol * 0.05


def test_force_is_one_fourier_mode(lib):
    """On a uniform state the right-hand side is zero, so after one step u equals the force:
    Re[(f_re + i f_im) e^{i (k.y + phase)}] sqrt(dt) cs magnitude sqrt(kaver cs). Its transform
    has weight only at +k and -k, with the complex amplitude that formula fixes."""
    n = (32, 32, 32)
    info, fp = forced_info(lib, n, (-1, 0, 0), relhel=0.4, magnitude=0.3)
    say(f"case forcing Fourier mode {lib.precision} k=({fp.k_force.x},{fp.k_force.y},{fp.k_force.z})")
    data = ac.HostMesh(lib, info).data.copy()
    data[:] = 0
    dt = 1e-3
    out = run_steps(lib, info, data, dt, 1)
    sl = O.interior_slices(out.shape[1:])
    for q in (0, 4, 5, 6, 7):
        assert np.all(out[q][sl] == 0.0)
    cs = np.sqrt(float(info["AC_cs2_sound"]))
    amp = np.sqrt(dt) * cs * float(fp.magnitude) * np.sqrt(float(fp.kaver) * cs)
    k = np.array([fp.k_force.x, fp.k_force.y, fp.k_force.z], dtype=float)
    fre = np.array([fp.ff_hel_re.x, fp.ff_hel_re.y, fp.ff_hel_re.z], dtype=float)
    fim = np.array([fp.ff_hel_im.x, fp.ff_hel_im.y, fp.ff_hel_im.z], dtype=float)
    N = n[0] * n[1] * n[2]
    # numpy index order (z, y, x); +k at (kz, ky, kx) mod n
    ik = tuple(int(round(v)) % m for v, m in zip(k[::-1], n[::-1]))
    imk = tuple((-int(round(v))) % m for v, m in zip(k[::-1], n[::-1]))
    tol = 1e-5 if lib.precision == "fp32" else 1e-12
    for d in range(3):
        u = out[1 + d][sl].astype(np.float64)
        F = np.fft.fftn(u)  # sum_x u(x) e^{-i m.y}
        # u = Re[c e^{i k.y}] with c = amp (f_re + i f_im) e^{i phase}: F[+k] = N c / 2
        c = amp * (fre[d] + 1j * fim[d]) * np.exp(1j * float(fp.phase))
        assert abs(F[ik] - N * c / 2) <= tol * N * amp, (d, F[ik], N * c / 2)
        assert abs(F[imk] - N * np.conj(c) / 2) <= tol * N * amp
        rest = np.abs(F).copy()
        rest[ik] = 0.0
        rest[imk] = 0.0
        assert rest.max() <= tol * N * amp
    # solenoidal: k . f = 0 for both parts
    assert abs(k @ fre) < 1e-6 * np.linalg.norm(k) and abs(k @ fim) < 1e-6 * np.linalg.norm(k)


@pytest.mark.parametrize("switch,forced", [(0.0, False), (-1e-9, True), (0.5, False)])
def test_switch_forcing(lib, switch, forced):
    """Forcing acts while AC_current_time > AC_switch_forcing (the time entering the step)."""
    say(f"case switch forcing {lib.precision} switch={switch}")
    n = (16, 16, 16)
    info, fp = forced_info(lib, n, (-1, 0, 0), AC_switch_forcing=switch, AC_current_time=0.0)
    off, _ = forced_info(lib, n, (-1, 0, 0), AC_lforcing=False)
    data = initial(lib, info)
    a = run_steps(lib, info, data, 1e-3, 1)
    b = run_steps(lib, off, data, 1e-3, 1)
    same = np.array_equal(a, b)
    assert same != forced


@pytest.mark.parametrize("n,shape", [((64, 16, 32), (64, 4, 16)), ((32, 32, 32), (-1, 0, 0))])
@pytest.mark.parametrize("mode", ["courant", "reference"])
def test_forced_multistep_matches_stepwise(lib, n, shape, mode):
    """acDeviceIntegrateStepsCFL with forcing (periodic images, device dt, forcing kernel doing
    the courant reduction) equals fill + acDeviceIntegrateSubstepDeviceDt + swap, bit for bit,
    and the next dt follows the forced state."""
    say(f"case forced multistep {lib.precision} n={n} shape={shape} mode={mode}")
    info, fp = forced_info(lib, n, shape)
    data = initial(lib, info)
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
    if mode == "courant":
        sl = O.interior_slices(outs[0][0].shape[1:])
        u2 = sum(outs[0][0][q][sl].astype(np.float64) ** 2 for q in (1, 2, 3))
        want = float(courant) * float(info["AC_dsmin"]) / (1.0 + np.sqrt(u2.max()))
        assert outs[0][1] == pytest.approx(want, rel=1e-12)


def test_multistep_switch_uses_device_time(lib):
    """Inside one acDeviceIntegrateStepsCFL call the forcing switches on once the accumulated
    simulated time passes AC_switch_forcing: same result as steps driven one at a time with
    AC_current_time advanced by each step's dt."""
    n = (32, 32, 32)
    courant = lib.np_real(0.3)
    info, fp = forced_info(lib, n, (-1, 0, 0), AC_current_time=0.0)
    data = initial(lib, info)
    # the dt of the first steps, to put the switch inside the third step's start times
    mesh = ac.HostMesh(lib, info, data)
    with ac.Device(lib, info) as dev:
        dev.load_mesh(mesh)
        dev.periodic()
        ac.check(lib.acDeviceUpdateTimestepCFL(dev.handle, 0, courant), "cfl")
        dt0 = ctypes.c_double()
        ac.check(lib.acDeviceGetTimestep(dev.handle, 0, ctypes.byref(dt0)), "dt")
    switch = 1.5 * dt0.value
    say(f"case multistep switch {lib.precision} dt0={dt0.value:.6g} switch={switch:.6g}")
    info["AC_switch_forcing"] = switch
    outs = []
    for whole in (True, False):
        mesh = ac.HostMesh(lib, info, data)
        with ac.Device(lib, info) as dev:
            dev.load_mesh(mesh)
            dev.periodic()
            ac.check(lib.acDeviceUpdateTimestepCFL(dev.handle, 0, courant), "cfl")
            if whole:
                t = ctypes.c_double(0.0)
                ac.check(lib.acDeviceIntegrateStepsCFL(dev.handle, 0, 4, courant, ctypes.byref(t)),
                         "steps")
            else:
                t_now = 0.0
                for _ in range(4):
                    dt = ctypes.c_double()
                    ac.check(lib.acDeviceGetTimestep(dev.handle, 0, ctypes.byref(dt)), "dt")
                    ac.check(lib.acDeviceLoadScalarUniform(dev.handle, 0, ac.AC_current_time,
                                                           lib.np_real(t_now)), "time")
                    ac.check(lib.acDeviceIntegrateStepsCFL(dev.handle, 0, 1, courant, None), "step")
                    t_now += dt.value
            dev.periodic()
            out = ac.HostMesh(lib, info)
            dev.store_mesh(out)
            outs.append(out.data.copy())
    assert np.array_equal(outs[0], outs[1])
