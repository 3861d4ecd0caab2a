"""Diagnostics timeseries in the reference's on-disk format.

The reference's standalone driver appends one row of reductions every `save_steps` steps to
"timeseries.ts" (ref samples/standalone_mpi/main.cc:218-391): a header line of column names
separated by two spaces, then rows "step t_step dt" followed by min / rms / max of |u|, |B|,
the Alfven speed and every field, printed with %e. `acHostAppendTimeseries` writes that format
from one `acDeviceReduceDiagnostics` pass; `read_timeseries` reads it back (also files the
reference wrote); `run_timeseries` is the driver loop: device-resident CFL steps with a row
every `every` steps, stopping at the first NaN as the reference driver does.
"""
import ctypes

import numpy as np

from . import astaroth as ac


def read_timeseries(path):
    """Returns (column names, rows as a float64 array of shape (nrows, ncols))."""
    with open(path) as fp:
        names = fp.readline().split()
        rows = [[float(t) for t in line.split()] for line in fp if line.strip()]
    data = np.array(rows, dtype=np.float64).reshape(len(rows), -1)
    if data.size and data.shape[1] != len(names):
        raise ValueError(f"{path}: {data.shape[1]} values per row, {len(names)} columns")
    return names, data


def append_row(lib, path, step, t_step, dt, diag):
    """acHostAppendTimeseries; returns True when a per-field reduction was NaN."""
    nan = ctypes.c_int(0)
    ac.check(lib.acHostAppendTimeseries(str(path).encode(), int(step), float(t_step), float(dt),
                                        ctypes.byref(diag), ctypes.byref(nan)),
             "acHostAppendTimeseries")
    return bool(nan.value)


def run_timeseries(dev, nsteps, every, path, courant=0.2, forcing_seed=312256655):
    """Integrates nsteps RK3 steps with the device CFL timestep, appending a diagnostics row
    at step 0 and every `every` steps. Returns (steps done, simulated time).

    With AC_lforcing set in the device's configuration, new helical forcing parameters are
    drawn before every AC_forcing_period_steps-th step and the step's start time is loaded as
    AC_current_time, as the reference driver's GenerateForcing action does (ref
    samples/standalone_mpi/main.cc:1377-1383,1447-1453,1567-1572; generator seeded once with
    `forcing_seed`, the driver's seed at main.cc:1302)."""
    lib = dev.lib
    info = dev.info
    forced = bool(info["AC_lforcing"])
    period = max(1, int(info["AC_forcing_period_steps"])) if forced else 0
    if forced:
        lib.seed_rng(forcing_seed)
    ac.check(lib.acDeviceUpdateTimestepCFL(dev.handle, ac.STREAM_DEFAULT, courant), "CFL")
    t = ctypes.c_double(0.0)
    step = 0
    while True:
        dt = ctypes.c_double()
        ac.check(lib.acDeviceGetTimestep(dev.handle, ac.STREAM_DEFAULT, ctypes.byref(dt)), "dt")
        dev.periodic()
        if append_row(lib, path, step, t.value, dt.value, dev.diagnostics()) or step >= nsteps:
            return step, t.value
        k = min(every, nsteps - step)
        while k > 0:
            chunk = k
            if forced:
                if step % period == 0:
                    fp = lib.generateHelicalForcingParams(info["AC_relhel"],
                                                          info["AC_forcing_magnitude"],
                                                          info["AC_kmin"], info["AC_kmax"])
                    ac.check(lib.acDeviceLoadForcingParams(dev.handle, ac.STREAM_DEFAULT, fp),
                             "acDeviceLoadForcingParams")
                ac.check(lib.acDeviceLoadScalarUniform(dev.handle, ac.STREAM_DEFAULT,
                                                       ac.AC_current_time, lib.np_real(t.value)),
                         "AC_current_time")
                chunk = min(k, period - step % period)
            ac.check(lib.acDeviceIntegrateStepsCFL(dev.handle, ac.STREAM_DEFAULT, chunk, courant,
                                                   ctypes.byref(t)), "acDeviceIntegrateStepsCFL")
            step += chunk
            k -= chunk
