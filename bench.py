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
    n = args.n
    n3 = n ** 3
    lib, dev = make_device(ac, "fp32", n, 0, tuple(args.shape))
    wall, solve_ms, launches = timed(lib, dev, args.steps, args.warmup)
    tx, ty, tz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.acDeviceGetSolveLaunchShape(dev.handle, ctypes.byref(tx), ctypes.byref(ty),
                                    ctypes.byref(tz))
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
                "solve_achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}
        out["variants"] = variants
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
