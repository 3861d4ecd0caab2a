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
