"""The diagnostics timeseries driver on the device: rows every `every` steps in the reference's
format, simulated time and dt consistent with the device CFL timestep."""
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac
from astaroth_amd import timeseries as ts

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_run_timeseries(prec, tmp_path, require_gpu):
    print(f"case timeseries {prec}", flush=True)
    sys.stdout.flush()
    lib = ac.load(prec)
    info = lib.config((16, 16, 16), AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0,
                      AC_nu_visc=5e-3, AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3)
    path = tmp_path / "timeseries.ts"
    with ac.Device(lib, info) as dev:
        dev.init_fourier(5, 16, 4, 0.05, 1e-4)
        steps, t = ts.run_timeseries(dev, 25, 10, path, courant=0.2)
    assert steps == 25
    names, data = ts.read_timeseries(path)
    assert list(data[:, 0]) == [0, 10, 20, 25]
    assert np.all(np.isfinite(data))
    # t_step accumulates the per-step dt; dt = 0.2 dx / (cs + max|u|) of each row's state
    dx = 2 * np.pi / 16
    umax = data[:, names.index("uu_total_max")]
    assert np.allclose(data[:, 2], 0.2 * dx / (1.0 + umax), rtol=1e-5)
    assert data[0, 1] == 0.0 and t == pytest.approx(data[-1, 1], rel=1e-6)
    assert np.all(np.diff(data[:, 1]) > 0)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_run_timeseries_forced(prec, tmp_path, require_gpu):
    """With AC_lforcing the driver draws new forcing parameters every AC_forcing_period_steps
    steps from the seeded generator and loads the step's start time: the same state as a
    loop doing that by hand one step at a time, and a state different from the unforced run."""
    print(f"case timeseries forced {prec}", flush=True)
    sys.stdout.flush()
    lib = ac.load(prec)
    common = dict(AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0, AC_nu_visc=5e-3,
                  AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3, AC_forcing_magnitude=0.05,
                  AC_relhel=0.5, AC_kmin=1.0, AC_kmax=3.0, AC_switch_forcing=-1.0)
    forced = lib.config((16, 16, 16), AC_lforcing=True, AC_forcing_period_steps=2, **common)
    plain = lib.config((16, 16, 16), **common)
    courant = lib.np_real(0.2)
    outs = []
    for info, driver in ((forced, True), (forced, False), (plain, True)):
        with ac.Device(lib, info) as dev:
            dev.init_fourier(5, 16, 4, 0.05, 1e-4)
            if driver:
                steps, t = ts.run_timeseries(dev, 7, 3, tmp_path / f"ts{len(outs)}.ts",
                                             courant=courant, forcing_seed=77)
                assert steps == 7
            else:
                lib.seed_rng(77)
                ac.check(lib.acDeviceUpdateTimestepCFL(dev.handle, 0, courant), "cfl")
                t_now = 0.0
                for step in range(7):
                    if step % 2 == 0:
                        fp = lib.generateHelicalForcingParams(lib.np_real(0.5), lib.np_real(0.05),
                                                              lib.np_real(1.0), lib.np_real(3.0))
                        ac.check(lib.acDeviceLoadForcingParams(dev.handle, 0, fp), "load")
                    ac.check(lib.acDeviceLoadScalarUniform(dev.handle, 0, ac.AC_current_time,
                                                           lib.np_real(t_now)), "time")
                    dt = ac.ctypes.c_double(0.0)
                    ac.check(lib.acDeviceIntegrateStepsCFL(dev.handle, 0, 1, courant,
                                                           ac.ctypes.byref(dt)), "step")
                    t_now += dt.value
            dev.periodic()
            out = ac.HostMesh(lib, info)
            dev.store_mesh(out)
            outs.append(out.data.copy())
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0], outs[2])
