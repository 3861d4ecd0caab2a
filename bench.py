#!/usr/bin/env python3
"""Benchmark of the MHD hot path on one MI355X: cell-updates/s of full RK3 steps at 256^3 FP32.

Workload (the plain method on GPU): 3D periodic compressible non-ideal MHD on [0, 2pi)^3 at
256^3 interior + 3-cell halo, 8 fields, 6th-order stencils, Williamson 2N RK3, cs2 = 1,
gamma = 5/3, nu = eta = chi = 5e-3, upwinding and entropy on; initial lnrho = ss = 0, uu and aa
each 16 seeded Fourier modes with |k| <= 4 and amplitude 0.05 (synthetic, generated in FP64).
dt = 0.2 dx / (cs + max|u|) is recomputed every step from an FP64 device reduction fused into
the last substep, so the step loop never synchronises with the host.

A step = periodic halo fill + fused solve/RK3 update, three times. The timed region holds
exactly `--steps` steps with the state resident in HBM, bracketed by device synchronisation.

Printed: one JSON line with the headline metric plus
  roofline      -- the solve kernel's algorithmic HBM bytes per launch / its mean launch time
                   (HIP events recorded by the library on the stream the kernel runs on),
                   against the 8 TB/s HBM peak; `traffic` from committed rocprofv3 counters;
  cpu_baseline  -- the FP64 two-pass OpenMP CPU restatement of the reference's CPU build
                   timed on this host on a bounded sample (rank 0, N = 1 only);
  variants      -- FP64 and Kahan-compensated FP32 cell-updates/s on the same GPU;
  adjoint       -- time of one adjoint RK3 step (three adjoint substeps with their halo fills)
                   and of one checkpointed reverse step (plus recomputing the step's first two
                   substeps and moving states), each against one forward step, at the same
                   grid and fixed dt.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
NF = 8


def algorithmic_bytes_per_cell_update(real_bytes, kahan_bits):
    """Compulsory HBM bytes per interior cell per full RK3 step (3 substeps).

    Per field: s=0 reads f, writes f' and w; s=1 reads f and w, writes f' and w; s=2 reads f and
    w, writes f' (w is not read at s=0 since alpha_0 = 0, and not stored at s=2, where it is
    dead): 10 array touches. Kahan compensation adds a read and a write per substep (6)."""
    b = 10 * NF * real_bytes
    if kahan_bits:
        b += 6 * NF * (kahan_bits // 8)
    return b


def per_launch_bytes(n3, real_bytes, kahan_bits):
    # mean over the three substeps of one step
    return n3 * algorithmic_bytes_per_cell_update(real_bytes, kahan_bits) / 3.0


def make_device(ac, prec, n, kahan_bits=0, shape=(0, 0, 0)):
    lib = ac.load(prec)
    info = lib.config((n, n, n), AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0,
                      AC_nu_visc=5e-3, AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3,
                      AC_lupwd=True, AC_lentropy=True, AC_kahan_bits=kahan_bits,
                      AC_solve_tile_x=shape[0], AC_solve_tile_y=shape[1],
                      AC_solve_zchunk=shape[2])
    mesh = ac.HostMesh(lib, info)
    ac.check(lib.acHostMeshInitFourier(ctypes.byref(mesh.struct), 2024, 16, 4, 0.05, 0.0),
             "acHostMeshInitFourier")
    dev = ac.Device(lib, info)
    dev.load_mesh(mesh)
    dev.sync()
    del mesh
    return lib, dev


def solve_kernel_name(lib, dev):
    """Which solve kernel the device's (autotuned or pinned) launch shape runs."""
    kind = ctypes.c_int()
    lib.acDeviceGetSolveKernel(dev.handle, ctypes.byref(kind))
    return {-1: "direct", 0: "delay-line", 1: "sweep"}.get(kind.value, str(kind.value))


def run_steps(lib, dev, steps, courant):
    ac_ok = lib.acDeviceIntegrateStepsCFL(dev.handle, 0, steps, courant, None)
    if ac_ok != 0:
        raise RuntimeError("acDeviceIntegrateStepsCFL failed")


def timed(lib, dev, steps, warmup, courant=0.2):
    lib.acDeviceUpdateTimestepCFL(dev.handle, 0, courant)
    run_steps(lib, dev, warmup, courant)
    lib.acDeviceSynchronizeStream(dev.handle, 32)
    lib.acDeviceSetSolveTiming(dev.handle, 1)
    lib.acDeviceSynchronizeStream(dev.handle, 32)
    t0 = time.perf_counter()
    run_steps(lib, dev, steps, courant)
    lib.acDeviceSynchronizeStream(dev.handle, 32)
    t1 = time.perf_counter()
    tot = ctypes.c_double()
    cnt = ctypes.c_int()
    lib.acDeviceGetSolveTiming(dev.handle, ctypes.byref(tot), ctypes.byref(cnt))
    lib.acDeviceSetSolveTiming(dev.handle, 0)
    return t1 - t0, tot.value, cnt.value


def reference_call_pattern(ac, n, steps=20, warmup=3, dt=1e-3):
    """The drop-in loop a caller of the reference API writes (ref samples/devicetest/main.cc:
    98-103): acDevicePeriodicBoundconds + acDeviceIntegrateSubstep + acDeviceSwapBuffers per
    substep, fixed dt, FP32, no fused multi-step path. Host-timed around device syncs."""
    lib, dev = make_device(ac, "fp32", n)
    try:
        dt = lib.np_real(dt)
        fixed_steps(dev, dt, warmup)
        wall = fixed_steps(dev, dt, steps)
    finally:
        dev.close()
    return {"cell_updates_per_s": n ** 3 * steps / wall, "steps": steps, "dt": float(dt),
            "calls": "acDevicePeriodicBoundconds + acDeviceIntegrateSubstep + "
                     "acDeviceSwapBuffers per substep"}


def cpu_baseline(n=96, min_seconds=10.0, max_steps=1000):
    """ref-cpu: the reference's CPU-build algorithm (two-pass RK3, FP64), restated in C++ with
    OpenMP over z on every host core. Bounded sample: n^3 cells, steps until min_seconds."""
    import oracle.oracle as O
    from astaroth_amd import astaroth as ac
    lib = ac.load("fp64")
    info = lib.config((n, n, n), AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0,
                      AC_nu_visc=5e-3, AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3)
    mesh = ac.HostMesh(lib, info)
    lib.acHostMeshInitFourier(ctypes.byref(mesh.struct), 2024, 16, 4, 0.05, 0.0)
    f = mesh.data.copy()
    w = np.zeros_like(f)
    cfg = O.config_from_info(info)
    threads = os.cpu_count() or 1
    O.refcpu_step(f, w, cfg, 1e-3, threads)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and (time.perf_counter() - t0) < min_seconds:
        O.refcpu_step(f, w, cfg, 1e-3, threads)
        steps += 1
    dt = time.perf_counter() - t0
    return dict(value=n ** 3 * steps / dt, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"ref-cpu FP64 two-pass RK3, {n}^3 interior, {steps} steps "
                       f"({dt:.1f} s) on {threads} host threads")


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes (a plain copy kernel of known traffic: nbytes read,
    nbytes written), so FETCH_SIZE / WRITE_SIZE counters can be scaled to bytes."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def load_traffic(path):
    try:
        with open(path) as fp:
            return json.load(fp)
    except (OSError, ValueError):
        return None


def adjoint_cost(ac, n, steps=5):
    """Wall time per step, device-synchronised, of forward steps, adjoint steps and full
    reverse steps of astaroth_amd.adjoint.MHDAdjoint on the benchmark state (FP32)."""
    import torch
    from astaroth_amd import adjoint as AD
    lib = ac.load("fp32")
    info = lib.config((n, n, n), AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0,
                      AC_nu_visc=5e-3, AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3,
                      AC_lupwd=True, AC_lentropy=True)
    mesh = ac.HostMesh(lib, info)
    ac.check(lib.acHostMeshInitFourier(ctypes.byref(mesh.struct), 2024, 16, 4, 0.05, 0.0),
             "acHostMeshInitFourier")
    dt = 1e-3
    runner = AD.MHDAdjoint(lib, info, steps, dt)
    try:
        runner.dev.load_mesh(mesh)
        del mesh
        runner._sync()
        lam_f = 1e-3 * torch.randn(runner.shape, dtype=runner.tdtype, device=runner.tdev)
        lam_w = torch.zeros_like(lam_f)
        torch.cuda.synchronize(runner.tdev)

        def clock(fn):
            fn()  # warm-up (first launches, autotuning)
            runner._sync()
            torch.cuda.synchronize(runner.tdev)
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            runner._sync()
            torch.cuda.synchronize(runner.tdev)
            return (time.perf_counter() - t0) / steps * 1e3

        def adjoint_step():
            for s in (2, 1, 0):
                runner.dev.periodic()
                ac.check(lib.acDeviceIntegrateSubstepAdjoint(
                    runner.dev.handle, 0, s, lib.np_real(dt), ctypes.c_void_p(lam_f.data_ptr()),
                    ctypes.c_void_p(lam_w.data_ptr()), ctypes.c_void_p(runner.workspace.data_ptr()),
                    ctypes.c_void_p(runner.param_grad.data_ptr())), "adjoint substep")

        fwd = clock(lambda: runner.advance(1))
        adj = clock(adjoint_step)
        rev = clock(lambda: runner.reverse_step(lam_f, lam_w))
        return {"grid": [n, n, n], "dtype": "f32", "steps": steps,
                "forward_ms_per_step": round(fwd, 3), "adjoint_ms_per_step": round(adj, 3),
                "reverse_ms_per_step": round(rev, 3),
                "adjoint_over_forward": round(adj / fwd, 3),
                "reverse_over_forward": round(rev / fwd, 3),
                "workspace_bytes": int(runner.workspace.numel())}
    finally:
        runner.close()


# ---------------------------------------------------------------------------------------------
# Measurement forms beyond the headline line (python bench.py --form NAME); each prints one JSON
# line. The GPU's own FP64 mode is the reference for the precision comparisons.

PHYS = dict(AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0, AC_nu_visc=5e-3,
            AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3, AC_lupwd=True, AC_lentropy=True)


def _torch():
    import torch
    return torch


def state_tensor(lib, dev):
    """The device state (halos filled) copied into a torch tensor on the same GPU."""
    torch = _torch()
    mx, my, mz = dev.info["AC_mlocal"]
    t = torch.empty((NF, mz, my, mx), dtype=torch.float32 if lib.precision == "fp32" else
                    torch.float64, device="cuda")
    dev.periodic()
    dev.sync()
    torch.cuda.synchronize()
    if lib.acDeviceStoreMeshToDevicePtr(dev.handle, 0, ctypes.c_void_p(t.data_ptr())):
        raise RuntimeError("acDeviceStoreMeshToDevicePtr failed")
    dev.sync()
    return t


def field_errors_gpu(ref, cand):
    """Per field: max |cand - ref| / max |ref| and the relative L2 error, over interior cells."""
    out = []
    for q in range(NF):
        m = ref[q, 3:-3, 3:-3, 3:-3].double()
        d = cand[q, 3:-3, 3:-3, 3:-3].double() - m
        scale = max(float(m.abs().max()), 1e-300)
        out.append({"maxrel": float(d.abs().max()) / scale,
                    "l2": float(d.pow(2).sum().sqrt()) / max(float(m.pow(2).sum().sqrt()), 1e-300)})
        del m, d
    return out


def fixed_steps(dev, dt, steps):
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.step(dt)
    dev.sync()
    return time.perf_counter() - t0


def form_precision(ac, n=512, steps=200, mid=100, seed=512, kbs=(8, 16, 32), nmodes=64, kmax=8,
                   ampl=0.1, perturb=1e-4, with_va=True):
    """512^3, 64 modes |k| <= 8, amplitude 0.1, lnrho + U(-1e-4, 1e-4); FP64 reference, plain FP32
    and Kahan-compensated FP32 (compensation stored in 8, 16 or 32 bits). This initial condition
    is magnetically dominated (Alfven speed well above cs + |u|), so dt is recomputed every step
    from the FP64 state as 0.2 dx / (cs + max|u| + max v_A) and the same dt sequence is replayed
    in every FP32 variant. Errors at `mid` and `steps` steps.

    Inputs are identical: the FP64 reference starts from the FP32-rounded initial condition the
    FP32 runs start from, so the errors measure the FP32 arithmetic and storage. A second FP64
    run from the unrounded condition measures what the rounding of the inputs alone grows into
    ("input_rounding")."""
    torch = _torch()
    n3 = n ** 3
    dx = 2 * np.pi / n
    l64 = ac.load("fp64")
    info64 = l64.config((n, n, n), **PHYS)
    dev = ac.Device(l64, info64)
    dev.init_fourier(seed, nmodes, kmax, ampl, perturb)
    exact_ic = state_tensor(l64, dev)
    rounded_ic = exact_ic.float().double().contiguous()
    torch.cuda.synchronize()
    if l64.acDeviceLoadMeshFromDevicePtr(dev.handle, 0, ctypes.c_void_p(rounded_ic.data_ptr())):
        raise RuntimeError("acDeviceLoadMeshFromDevicePtr failed")
    dev.sync()
    del rounded_ic
    dts, vmax = [], []
    t_ref = 0.0
    refs = {}
    for k in range(steps):
        dev.periodic()
        d = dev.diagnostics()
        dt = 0.2 * dx / (1.0 + d.uu_max + (d.va_max if with_va else 0.0))
        dts.append(dt)
        vmax.append((d.uu_max, d.va_max))
        dev.sync()
        t0 = time.perf_counter()
        dev.step(l64.np_real(dt))
        dev.sync()
        t_ref += time.perf_counter() - t0
        if k + 1 in (mid, steps):
            refs[k + 1] = state_tensor(l64, dev)
    dev.close()
    runs = {"fp64": {"cell_updates_per_s": n3 * steps / t_ref}}
    # the FP64 run from the unrounded initial condition, same dt sequence
    dev = ac.Device(l64, info64)
    torch.cuda.synchronize()
    if l64.acDeviceLoadMeshFromDevicePtr(dev.handle, 0, ctypes.c_void_p(exact_ic.data_ptr())):
        raise RuntimeError("acDeviceLoadMeshFromDevicePtr failed")
    dev.sync()
    del exact_ic
    rec = {}
    for k in range(steps):
        dev.step(l64.np_real(dts[k]))
        if k + 1 in refs:
            e = field_errors_gpu(refs[k + 1], state_tensor(l64, dev))
            rec[f"errors_after_{k + 1}"] = {ac.FIELD_NAMES[q]: e[q] for q in range(NF)}
            rec[f"max_maxrel_after_{k + 1}"] = max(x["maxrel"] for x in e)
    dev.close()
    torch.cuda.empty_cache()
    runs["fp64_unrounded_ic"] = rec
    for kb in (0,) + tuple(kbs):
        l32 = ac.load("fp32")
        dv = ac.Device(l32, l32.config((n, n, n), AC_kahan_bits=kb, **PHYS))
        dv.init_fourier(seed, nmodes, kmax, ampl, perturb)
        t_run = 0.0
        errs = {}
        for k in range(steps):
            dv.sync()
            t0 = time.perf_counter()
            dv.step(l32.np_real(dts[k]))
            dv.sync()
            t_run += time.perf_counter() - t0
            if k + 1 in refs:
                errs[k + 1] = field_errors_gpu(refs[k + 1], state_tensor(l32, dv))
        dv.close()
        torch.cuda.empty_cache()
        rec = {"cell_updates_per_s": n3 * steps / t_run}
        for k, e in errs.items():
            rec[f"errors_after_{k}"] = {ac.FIELD_NAMES[q]: e[q] for q in range(NF)}
            rec[f"max_maxrel_after_{k}"] = max(x["maxrel"] for x in e)
        runs["fp32" if kb == 0 else f"fp32_kahan{kb}"] = rec
    return {"form": "precision", "grid": [n, n, n], "steps": steps,
            "dt_first_last": [dts[0], dts[-1]], "max_u_first_last": [vmax[0][0], vmax[-1][0]],
            "max_vA_first_last": [vmax[0][1], vmax[-1][1]],
            "ic": f"{nmodes} modes |k|<={kmax} ampl {ampl}, lnrho U(-{perturb},{perturb}), seed {seed}",
            "dt_rule": "0.2 dx / (cs + max|u|" + (" + max vA)" if with_va else ")") +
                       " of the FP64 state, every step, replayed in the FP32 runs",
            "timing": "wall time of the steps only, device-synchronised per step",
            "reference": "FP64 GPU mode from the FP32-rounded initial condition (identical "
                         "inputs); fp64_unrounded_ic: FP64 from the unrounded one",
            "runs": runs}


def form_scale(ac, n=1024, steps=50, kb=16, seed=1024):
    """1024^3 FP32 with Kahan compensation (kb-bit), 256 modes |k| <= 16, amplitude 0.1, device
    CFL dt, autotuned launch shape, initial condition generated on the device."""
    lib = ac.load("fp32")
    info = lib.config((n, n, n), AC_kahan_bits=kb, **PHYS)
    t0 = time.perf_counter()
    dev = ac.Device(lib, info)
    dev.init_fourier(seed, 256, 16, 0.1, 0.0)
    t_ic = time.perf_counter() - t0
    wall, solve_ms, launches = timed(lib, dev, steps, 3)
    tx, ty, tz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.acDeviceGetSolveLaunchShape(dev.handle, ctypes.byref(tx), ctypes.byref(ty), ctypes.byref(tz))
    dev.close()
    m3 = (n + 6) ** 3
    resident = m3 * NF * (3 * 4 + kb // 8)
    bl = per_launch_bytes(n ** 3, 4, kb)
    ach = bl / (solve_ms / 1e3 / max(launches, 1)) / 1e9
    return {"form": "scale", "grid": [n, n, n], "dtype": "f32", "kahan_bits": kb, "steps": steps,
            "cell_updates_per_s": n ** 3 * steps / wall, "ms_per_step": wall / steps * 1e3,
            "resident_bytes": resident, "ic_seconds": t_ic,
            "launch_shape": [tx.value, ty.value, tz.value],
            "solve_achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}


def form_reductions(ac, n=256, reps=50, seed=2024):
    """The on-device reductions at 256^3 FP32 (the vrms / max|u| / per-field min-max path):
    acDeviceReduceScal (RTYPE_MAX, RTYPE_RMS), acDeviceReduceVec (RTYPE_MAX), acDeviceReduceVecScal
    (RTYPE_ALFVEN_MAX) and the one-pass acDeviceReduceDiagnostics (every field plus |u|, |curl A|,
    v_A). Host-timed per call (each call synchronises: it returns the value to the host); the
    algorithmic bytes are the interior values read (4 B each)."""
    lib, dev = make_device(ac, "fp32", n)
    try:
        dev.periodic()
        dev.sync()
        n3 = n ** 3
        calls = {
            "scal_max": (lambda: dev.reduce_scal("RTYPE_MAX", 1), 1),
            "scal_rms": (lambda: dev.reduce_scal("RTYPE_RMS", 1), 1),
            "vec_max": (lambda: dev.reduce_vec("RTYPE_MAX", 1, 2, 3), 3),
            "vecscal_alfven_max": (lambda: dev.reduce_vec_scal("RTYPE_ALFVEN_MAX", 4, 5, 6, 0), 4),
            "diagnostics": (lambda: dev.diagnostics(), NF),
        }
        out = {"form": "reductions", "grid": [n, n, n], "dtype": "f32", "reps": reps}
        for name, (fn, nbuf) in calls.items():
            fn()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            t = (time.perf_counter() - t0) / reps
            out[name] = {"us_per_call": round(t * 1e6, 1),
                         "GBs_algorithmic": round(nbuf * n3 * 4 / t / 1e9, 1)}
        return out
    finally:
        dev.close()


def form_scale_adjoint(ac, n=384, steps=500, snaps=24, seed=384):
    """Adjoint of a 384^3 FP32 run of 500 fixed-dt steps with revolve checkpointing (snapshot
    slots sized to the HBM budget), loss 0.5 |u(T) - u_target|^2 with u_target the initial u of a
    second seed; reports forward, reverse and recomputation cost."""
    torch = _torch()
    from astaroth_amd import adjoint as AD
    lib = ac.load("fp32")
    info = lib.config((n, n, n), **PHYS)
    runner = AD.MHDAdjoint(lib, info, steps, 1.0, schedule="revolve", snaps=snaps)
    try:
        runner.dev.init_fourier(seed, 256, 16, 0.1, 0.0)
        runner.dev.periodic()
        d = runner.dev.diagnostics()
        # fixed dt for the whole run: the field is strong (|k| up to 16), so the Alfven speed
        # enters the CFL bound; a fifth of it leaves room for the flow to speed up
        runner.dt = lib.np_real(0.2 * 0.2 * (2 * np.pi / n) / (1.0 + d.uu_max + d.va_max))
        f0 = state_tensor(lib, runner.dev)
        runner.dev.init_fourier(seed + 7, 256, 16, 0.1, 0.0)
        tgt = state_tensor(lib, runner.dev)[1:4].clone()
        f0.requires_grad_(True)
        nu = torch.tensor(5e-3, dtype=torch.float32, requires_grad=True)
        eta = torch.tensor(5e-3, dtype=torch.float32, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = AD.MHDSolve.apply(f0, nu, eta, runner)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        fwd_steps = runner.forward_steps
        loss = 0.5 * ((out[1:4, 3:-3, 3:-3, 3:-3] - tgt[:, 3:-3, 3:-3, 3:-3]).double() ** 2).sum()
        t0 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        rec = runner.forward_steps - fwd_steps
        return {"form": "scale_adjoint", "grid": [n, n, n], "dtype": "f32", "steps": steps,
                "schedule": f"revolve, {snaps} snapshot slots of {NF * (n + 6) ** 3 * 4} bytes",
                "forward_s": t_fwd, "reverse_s": t_bwd,
                "reverse_over_forward": t_bwd / t_fwd,
                "recomputed_steps": rec, "recomputed_per_step": rec / steps,
                "dt": float(runner.dt), "loss": float(loss.detach()),
                "grad_nu": float(nu.grad), "grad_eta": float(eta.grad)}
    finally:
        runner.close()


def form_differentiation(ac, n=128, steps=50, every=10, seed=128, ndir=8):
    """128^3, 50 fixed-dt steps, checkpoints every 10 steps; loss 0.5 |u(T) - u_target|^2 with
    u_target the state of a second seeded run. FP32 adjoint gradient, FP64 adjoint gradient, and
    centred differences of the FP64 loss (h = 1e-3 without upwinding, where the loss is smooth;
    h = 1e-5 with it) on ndir random directions plus nu and eta."""
    torch = _torch()
    from astaroth_amd import adjoint as AD
    res = {"form": "differentiation", "grid": [n, n, n], "steps": steps,
           "checkpoint_every": every}
    dx = 2 * np.pi / n
    dt = 0.2 * dx / 1.2
    res["dt"] = dt
    gen = torch.Generator(device="cpu").manual_seed(seed)
    for upwd in (False, True):
        out = {}
        grads = {}
        for prec in ("fp64", "fp32"):
            lib = ac.load(prec)
            info = lib.config((n, n, n), **dict(PHYS, AC_lupwd=upwd))
            runner = AD.MHDAdjoint(lib, info, steps, dt, schedule="uniform", every=every)
            try:
                runner.dev.init_fourier(seed + 1, 16, 4, 0.05, 1e-4)
                runner.advance(steps)
                tgt = state_tensor(lib, runner.dev)[1:4].clone()
                runner.dev.init_fourier(seed, 16, 4, 0.05, 1e-4)
                f0 = state_tensor(lib, runner.dev)
                ft = f0.clone().requires_grad_(True)
                nu = torch.tensor(5e-3, dtype=runner.tdtype, requires_grad=True)
                eta = torch.tensor(5e-3, dtype=runner.tdtype, requires_grad=True)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                y = AD.MHDSolve.apply(ft, nu, eta, runner)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                L = 0.5 * ((y[1:4, 3:-3, 3:-3, 3:-3] - tgt[:, 3:-3, 3:-3, 3:-3]).double() ** 2).sum()
                L.backward()
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                grads[prec] = (ft.grad.double(), float(nu.grad), float(eta.grad))
                out[prec] = {"forward_s": t1 - t0, "backward_s": t2 - t1,
                             "backward_over_forward": (t2 - t1) / (t1 - t0)}
                if prec == "fp64":
                    f0_64, tgt_64, runner64 = f0, tgt, runner
                    continue
            finally:
                if prec != "fp64":
                    runner.close()

        def loss64(f, nu_v, eta_v):
            y, _ = runner64.run_forward(f, nu_v, eta_v)
            return 0.5 * float(((y[1:4, 3:-3, 3:-3, 3:-3] - tgt_64[:, 3:-3, 3:-3, 3:-3]) ** 2).sum())

        h = 1e-5 if upwd else 1e-3
        rows = []
        for k in range(ndir):
            v = torch.zeros_like(f0_64)
            v[:, 3:-3, 3:-3, 3:-3] = 0.01 * torch.randn(v[:, 3:-3, 3:-3, 3:-3].shape,
                                                        generator=gen, dtype=torch.float64).cuda()
            fd = (loss64(f0_64 + h * v, 5e-3, 5e-3) - loss64(f0_64 - h * v, 5e-3, 5e-3)) / (2 * h)
            a64 = float((grads["fp64"][0] * v).sum())
            a32 = float((grads["fp32"][0] * v).sum())
            rows.append({"fd": fd, "rel_fp64": abs(a64 - fd) / abs(fd),
                         "rel_fp32": abs(a32 - fd) / abs(fd)})
        for name, idx in (("nu", 1), ("eta", 2)):
            hp = 1e-3 * 5e-3
            a = (5e-3 + hp, 5e-3) if name == "nu" else (5e-3, 5e-3 + hp)
            b = (5e-3 - hp, 5e-3) if name == "nu" else (5e-3, 5e-3 - hp)
            fd = (loss64(f0_64, *a) - loss64(f0_64, *b)) / (2 * hp)
            rows.append({"param": name, "fd": fd,
                         "rel_fp64": abs(grads["fp64"][idx] - fd) / abs(fd),
                         "rel_fp32": abs(grads["fp32"][idx] - fd) / abs(fd)})
        runner64.close()
        out["fd_step"] = h
        out["checks"] = rows
        out["max_rel_fp64"] = max(r["rel_fp64"] for r in rows)
        out["max_rel_fp32"] = max(r["rel_fp32"] for r in rows)
        res["upwind" if upwd else "no_upwind"] = out
        torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--variant-steps", type=int, default=30)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--form", choices=["precision", "precision_plain", "scale", "scale_adjoint",
                                       "differentiation", "reductions"],
                    help="run one measurement form instead of the headline benchmark")
    ap.add_argument("--shape", type=int, nargs=3, default=(0, 0, 0),
                    help="pin the solve launch shape TX TY ZC (0 0 0 = autotune)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="after the run, copy this many bytes device-to-device (counter scale)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the single-device "
              "hot path has no multi-GPU mode in this build.", file=sys.stderr)
        sys.exit(2)

    from astaroth_amd import astaroth as ac
    if args.form:
        fn = {"precision": form_precision, "scale": form_scale, "reductions": form_reductions,
              "scale_adjoint": form_scale_adjoint, "differentiation": form_differentiation,
              # the plain-method workload (256^3, 16 modes |k| <= 4, amplitude 0.05, 100 steps,
              # dt = 0.2 dx / (cs + max|u|)) in the three precisions
              "precision_plain": lambda ac_: form_precision(ac_, n=256, steps=100, mid=50,
                                                            seed=2024, nmodes=16, kmax=4,
                                                            ampl=0.05, perturb=0.0, with_va=False)}
        print(json.dumps(fn[args.form](ac)), flush=True)
        return
    n = args.n
    n3 = n ** 3
    lib, dev = make_device(ac, "fp32", n, 0, tuple(args.shape))
    wall, solve_ms, launches = timed(lib, dev, args.steps, args.warmup)
    tx, ty, tz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.acDeviceGetSolveLaunchShape(dev.handle, ctypes.byref(tx), ctypes.byref(ty),
                                    ctypes.byref(tz))
    kernel_kind = solve_kernel_name(lib, dev)
    dev.close()
    value = n3 * args.steps / wall
    bytes_launch = per_launch_bytes(n3, 4, 0)
    mean_launch_s = solve_ms / 1e3 / max(launches, 1)
    achieved = bytes_launch / mean_launch_s / 1e9
    traffic_rec = load_traffic(os.path.join(REPO, "profiles", f"traffic_solve_{n}_fp32.json"))
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_rec.get("bytes_per_launch") if traffic_rec else None,
        "kernel": "solve (fused RK3 substep)",
        "algorithmic_bytes_per_launch": bytes_launch,
        "bytes_per_cell_update": algorithmic_bytes_per_cell_update(4, 0),
        "mean_launch_ms": round(mean_launch_s * 1e3, 4),
        "launches": launches,
        "launch_shape": [tx.value, ty.value, tz.value],
        "launch_kernel": kernel_kind,
    }
    out = {
        "metric": "MHD cell-updates/s at 256^3 fp32",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Fourier-mode initial condition)",
        "config": {"workload": f"mhd_rk3_{n}^3_fp32_cfl_dt", "grid": [n, n, n], "fields": 8,
                   "stencil_order": 6, "upwind": True, "entropy": True,
                   "dt": "0.2*dx/(cs+max|u|) per step, FP64 device reduction",
                   "parallelism": "single device"},
        "roofline": roofline,
    }
    if not args.no_variants:
        variants = {}
        for prec, kb in (("fp64", 0), ("fp32", 8), ("fp32", 16), ("fp32", 32)):
            try:
                lib_v, dev_v = make_device(ac, prec, n, kb)
                w_v, s_ms, cnt = timed(lib_v, dev_v, args.variant_steps, 3)
                shp = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int())
                lib_v.acDeviceGetSolveLaunchShape(dev_v.handle, *[ctypes.byref(v) for v in shp])
                kind_v = solve_kernel_name(lib_v, dev_v)
                dev_v.close()
            except Exception as e:  # report, do not hide
                variants[f"{prec}_kahan{kb}"] = {"error": str(e)}
                continue
            rb = 8 if prec == "fp64" else 4
            bl = per_launch_bytes(n3, rb, kb)
            ach = bl / (s_ms / 1e3 / max(cnt, 1)) / 1e9
            variants[f"{prec}_kahan{kb}"] = {
                "cell_updates_per_s": n3 * args.variant_steps / w_v,
                "bytes_per_cell_update": algorithmic_bytes_per_cell_update(rb, kb),
                "solve_achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                "launch_shape": [v.value for v in shp], "launch_kernel": kind_v}
        out["variants"] = variants
        try:
            out["reference_call_pattern"] = reference_call_pattern(ac, n)
        except Exception as e:  # report, do not hide
            out["reference_call_pattern"] = {"error": str(e)}
        fp64 = variants.get("fp64_kahan0", {}).get("cell_updates_per_s")
        if fp64:
            out["fp32_over_fp64"] = value / fp64
            for kb in (8, 16, 32):
                v = variants.get(f"fp32_kahan{kb}", {}).get("cell_updates_per_s")
                if v:
                    out[f"fp32_kahan{kb}_over_fp64"] = v / fp64
    if not args.no_adjoint:
        try:
            out["adjoint"] = adjoint_cost(ac, n)
        except Exception as e:  # report, do not hide
            out["adjoint"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
