"""acDeviceInitFourier (the seeded Fourier initial condition generated on the device) against
acHostMeshInitFourier followed by a load: identical bits, halos included."""
import ctypes
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("n,nmodes,kmax,ampl,pert", [((32, 32, 32), 16, 4, 0.05, 0.0),
                                                     ((24, 16, 20), 64, 8, 0.1, 1e-4),
                                                     ((16, 16, 16), 0, 1, 0.1, 1e-4)])
def test_device_ic_matches_host(prec, n, nmodes, kmax, ampl, pert, require_gpu):
    say(f"case device IC {prec} n={n} modes={nmodes} kmax={kmax} perturb={pert}")
    lib = ac.load(prec)
    info = lib.config(n, AC_len=(2 * np.pi,) * 3)
    host = ac.HostMesh(lib, info)
    ac.check(lib.acHostMeshInitFourier(ctypes.byref(host.struct), 77, nmodes, kmax, ampl, pert),
             "acHostMeshInitFourier")
    with ac.Device(lib, info) as dev:
        dev.init_fourier(77, nmodes, kmax, ampl, pert)
        out = ac.HostMesh(lib, info)
        dev.store_mesh(out)
    diff = np.argwhere(out.data != host.data)
    say(f"  differing values: {len(diff)}")
    assert len(diff) == 0, diff[:5]
    if nmodes:
        assert np.abs(out.data[1:7]).max() > 0.0
    if pert:
        assert 0.0 < np.abs(out.data[0]).max() <= pert
