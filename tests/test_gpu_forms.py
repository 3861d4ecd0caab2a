"""Checks at the full sizes of the precision and scale configurations, where the oracle is too
slow to run; each uses a closed form that holds at any N or an invariant of the adjoint:

  * 512^3 (the precision form's size) with 16- and 32-bit Kahan compensation and in FP64, and
    1024^3 with 16-bit compensation (the scale form): a Beltrami A-field at rest decays by exactly
    P3(-eta k2 dt) per step (k2 the 6th-order second-derivative symbol), and a shear wave by
    P3(-nu k2 dt); both through the multi-step device-dt path, whose solve reads the periodic
    images itself;
  * the plain-method precision form at 128^3: error bounds of plain, 8-bit and precise FP32
    after 100 steps against the FP64 mode;
  * 384^3 (the scale form's adjoint size), revolve checkpointing over several steps: the FP64
    gradient of 0.5 |u(T) - u_target|^2 against centred differences of the loss, the gradient
    with a different number of snapshot slots (bit for bit), and the FP32 gradient against the
    FP64 one.
The states are built and compared on the GPU with torch (device memory only; every step runs in
the library's HIP kernels through the C ABI)."""
import ctypes
import os
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac

pytestmark = pytest.mark.gpu

NF = ac.NUM_VTXBUF_HANDLES


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def p3(z):
    return 1.0 + z + z * z / 2.0 + z * z * z / 6.0


def k2_symbol(k, dx):
    t = k * dx
    return (49.0 / 18.0 - 3.0 * np.cos(t) + 0.3 * np.cos(2 * t) - np.cos(3 * t) / 45.0) / dx ** 2


def info_for(lib, n, **over):
    p = dict(AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0, AC_nu_visc=5e-3,
             AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3, AC_lupwd=True, AC_lentropy=True)
    p.update(over)
    return lib.config((n, n, n), **p)


def torch_dtype(lib):
    import torch
    return torch.float32 if lib.precision == "fp32" else torch.float64


def plane_wave_state(lib, n, fields):
    """(NF, m, m, m) device tensor, zero except fields[q] = a * fn(k x) along x (halos included:
    x of array index i is (i - 3) dx)."""
    import torch
    m = n + 6
    dx = 2 * np.pi / n
    x = (torch.arange(m, dtype=torch.float64, device="cuda") - 3) * dx
    t = torch.zeros((NF, m, m, m), dtype=torch_dtype(lib), device="cuda")
    for q, (a, k, fn) in fields.items():
        t[q] = (a * fn(k * x)).to(t.dtype)[None, None, :]
    return t


def run_multistep(lib, info, state, steps, courant=0.2):
    """Loads the device tensor, runs `steps` device-dt steps and returns (state, total time)."""
    import torch
    dev = ac.Device(lib, info)
    try:
        torch.cuda.synchronize()
        ac.check(lib.acDeviceLoadMeshFromDevicePtr(dev.handle, 0, ctypes.c_void_p(state.data_ptr())),
                 "load")
        dev.periodic()
        ac.check(lib.acDeviceUpdateTimestepCFL(dev.handle, 0, lib.np_real(courant)), "cfl")
        t = ctypes.c_double(0.0)
        ac.check(lib.acDeviceIntegrateStepsCFL(dev.handle, 0, steps, lib.np_real(courant),
                                               ctypes.byref(t)), "steps")
        dev.periodic()
        dev.sync()
        ac.check(lib.acDeviceStoreMeshToDevicePtr(dev.handle, 0, ctypes.c_void_p(state.data_ptr())),
                 "store")
        dev.sync()
    finally:
        dev.close()
    return state, t.value


def interior(t):
    return t[..., 3:-3, 3:-3, 3:-3]


BELTRAMI_TOL = {"fp32": 2e-6, "fp64": 1e-11}


def beltrami_check(lib, n, kb, steps=3):
    """A = a (0, sin kx, cos kx): j x B = 0 exactly, so A decays by P3(-eta k2 dt) per step and
    u, lnrho, A_x stay zero (up to rounding)."""
    import torch
    k, a = 3, 0.1
    info = info_for(lib, n, AC_kahan_bits=kb)
    state = plane_wave_state(lib, n, {5: (a, k, torch.sin), 6: (a, k, torch.cos)})
    out, total = run_multistep(lib, info, state, steps)
    dx = 2 * np.pi / n
    g = p3(-5e-3 * k2_symbol(k, dx) * total / steps) ** steps
    x = (torch.arange(n, dtype=torch.float64, device="cuda")) * dx
    errs = []
    for q, fn in ((5, torch.sin), (6, torch.cos)):
        ref = (a * g * fn(k * x))[None, None, :]
        errs.append(float((interior(out[q]).double() - ref).abs().max()) / a)
    rest = max(float(interior(out[q]).abs().max()) for q in (0, 1, 2, 3, 4))
    say(f"  Beltrami {lib.precision} kb={kb} n={n}: decay {g:.12f} max err / a = {max(errs):.3g}"
        f"  max |u, lnrho, A_x| = {rest:.3g}")
    assert max(errs) < BELTRAMI_TOL[lib.precision]
    # u stays at the rounding scale of j x B (see test_gpu_fullsize); far below the decay
    assert rest < (1e-6 if lib.precision == "fp32" else 1e-13)
    assert 1.0 - g > 1e-5
    del out, state
    torch.cuda.empty_cache()


def shear_check(lib, n, kb, steps=3):
    """u = (0, a sin kx, 0), entropy off: viscous decay by P3(-nu k2 dt) per step."""
    import torch
    k, a = 2, 0.05
    info = info_for(lib, n, AC_kahan_bits=kb, AC_lentropy=False)
    state = plane_wave_state(lib, n, {2: (a, k, torch.sin)})
    out, total = run_multistep(lib, info, state, steps)
    dx = 2 * np.pi / n
    g = p3(-5e-3 * k2_symbol(k, dx) * total / steps) ** steps
    x = (torch.arange(n, dtype=torch.float64, device="cuda")) * dx
    err = float((interior(out[2]).double() - (a * g * torch.sin(k * x))[None, None, :]).abs().max()) / a
    rest = max(float(interior(out[q]).abs().max()) for q in (0, 1, 3, 4, 5, 6))
    say(f"  shear {lib.precision} kb={kb} n={n}: decay {g:.12f} max err / a = {err:.3g}"
        f"  max other = {rest:.3g}")
    assert err < BELTRAMI_TOL[lib.precision]
    assert rest < 1e-6 * a
    assert 1.0 - g > 1e-6
    del out, state
    torch.cuda.empty_cache()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec,kb", [("fp32", 16), ("fp32", 32), ("fp64", 0)])
def test_precision_form_512_closed_forms(prec, kb, require_gpu):
    """The precision form's 512^3 legs (16- and 32-bit compensated FP32, FP64)."""
    lib = ac.load(prec)
    say(f"case 512^3 closed forms {prec} kb={kb}")
    beltrami_check(lib, 512, kb)
    shear_check(lib, 512, kb)


@pytest.mark.timeout(600)
def test_scale_form_1024_kahan16(require_gpu):
    """The scale form: 1024^3 FP32 with 16-bit compensation (about 122 GB of state resident)."""
    lib = ac.load("fp32")
    say("case 1024^3 closed form fp32 kb=16")
    beltrami_check(lib, 1024, 16, steps=2)


@pytest.mark.timeout(300)
def test_precision_form_128_bounds(require_gpu):
    """The plain-method precision form (16 modes |k| <= 4, amplitude 0.05, 100 steps, dt = 0.2 dx /
    (cs + max|u|) from the FP64 state) at 128^3: max-rel error per field after 100 steps against
    the FP64 mode started from the same FP32-rounded state. The precise mode (32-bit records,
    cancellation-free stencils) holds u and A below 1e-6 (measured u 6.4-6.9e-7, A 3.7-4.1e-7,
    lnrho 1.07e-6, ss 6.8e-6) and cuts the u, A and lnrho errors of plain FP32 (u 3.7-4.4e-6,
    A 1.2-1.8e-6, lnrho 4.5e-6) at least two-fold; the 8-bit record also beats plain FP32 in u
    (2.3-2.9e-6). The bounds leave about 45% over the measured values for a different launch
    shape's rounding."""
    import bench
    say("case precision form 128^3, 100 steps, plain / kahan8 / kahan32 against fp64")
    r = bench.form_precision(ac, n=128, steps=100, mid=50, seed=2024, kbs=(8, 32), nmodes=16,
                             kmax=4, ampl=0.05, perturb=0.0, with_va=False)
    groups = {"lnrho": ["VTXBUF_LNRHO"], "u": ["VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ"],
              "A": ["VTXBUF_AX", "VTXBUF_AY", "VTXBUF_AZ"], "ss": ["VTXBUF_ENTROPY"]}
    err = {run: {g: max(rec["errors_after_100"][q]["maxrel"] for q in qs)
                 for g, qs in groups.items()}
           for run, rec in r["runs"].items() if "errors_after_100" in rec}
    for run, e in err.items():
        say(f"  {run}: " + " ".join(f"{g} {v:.3g}" for g, v in e.items()))
    # the input rounding alone (FP64 from the unrounded state) stays far below every FP32 error
    assert max(err["fp64_unrounded_ic"].values()) < 1e-6
    precise, plain = err["fp32_kahan32"], err["fp32"]
    assert precise["u"] < 1e-6 and precise["A"] < 6e-7
    assert precise["lnrho"] < 1.6e-6 and precise["ss"] < 1e-5
    for g in ("u", "A", "lnrho"):
        assert precise[g] < 0.5 * plain[g]
    assert err["fp32_kahan8"]["u"] < plain["u"]


def _adjoint_modules():
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import test_gpu_adjoint as TA
    from astaroth_amd import adjoint as AD
    return TA, AD


@pytest.mark.timeout(900)
def test_scale_adjoint_384_revolve(require_gpu):
    """The scale form's adjoint size, 384^3, 4 steps reversed with revolve on 1 and on 2 snapshot
    slots: FP64 gradient against centred differences of the loss along a random direction and
    along nu, the two schedules bit for bit, and the FP32 gradient within 1e-4 of the FP64 one."""
    import torch
    TA, AD = _adjoint_modules()
    n, nsteps, dt = 384, 4, 2e-3
    grads = {}
    info64, f0_64, tgt_64 = TA.loss_and_target(ac.load("fp64"), n)
    # one pinned solve kernel, so that every runner computes the forward states alike
    pin = dict(AC_solve_tile_x=64, AC_solve_tile_y=4, AC_solve_zchunk=128, AC_solve_kernel=0)
    for prec in ("fp64", "fp32"):
        lib = ac.load(prec)
        if prec == "fp64":
            f0, tgt = f0_64, tgt_64
        else:  # the same initial state and target, rounded to FP32
            f0, tgt = f0_64.astype(np.float32), tgt_64.astype(np.float32)
        info = TA.physics_info(lib, n, AC_lupwd=True, **pin)
        for snaps in ((1, 2) if prec == "fp64" else (2,)):
            runner = AD.MHDAdjoint(lib, info, nsteps, dt, schedule="revolve", snaps=snaps)
            try:
                g, gnu, geta = TA.adjoint_gradient(runner, f0, tgt, 5e-3, 5e-3)
                say(f"case 384^3 revolve {prec} snaps={snaps}: forward steps {runner.forward_steps}")
                grads[(prec, snaps)] = (g, gnu, geta)
                if prec == "fp64" and snaps == 2:
                    h = TA.FD_STEP[True]
                    v = TA.random_direction(f0, 21)
                    fd = (TA.run_loss(runner, f0 + h * v, tgt, 5e-3, 5e-3) -
                          TA.run_loss(runner, f0 - h * v, tgt, 5e-3, 5e-3)) / (2 * h)
                    ad = float((g.cpu().numpy() * v).sum())
                    rel = abs(fd - ad) / abs(fd)
                    say(f"  direction: FD {fd:.12g} adjoint {ad:.12g} rel {rel:.3g}")
                    assert rel < 1e-6
                    del v
                    hp = 1e-3 * 5e-3
                    fd = (TA.run_loss(runner, f0, tgt, 5e-3 + hp, 5e-3) -
                          TA.run_loss(runner, f0, tgt, 5e-3 - hp, 5e-3)) / (2 * hp)
                    rel = abs(fd - gnu) / abs(fd)
                    say(f"  d/dnu: FD {fd:.12g} adjoint {gnu:.12g} rel {rel:.3g}")
                    assert rel < 1e-6
            finally:
                runner.close()
    f0_dir_shape = np.zeros(f0_64.shape, dtype=np.float64)
    del f0_64, tgt_64, f0, tgt
    a, b = grads[("fp64", 1)], grads[("fp64", 2)]
    assert torch.equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]
    g64 = grads[("fp64", 2)][0]
    g32 = grads[("fp32", 2)][0].double()
    d = g64 - g32
    rel_max = float(d.abs().max() / g64.abs().max())
    rel_l2 = float(d.norm() / g64.norm())
    v = torch.from_numpy(TA.random_direction(f0_dir_shape, 33)).to(g64.device)
    rel_dir = abs(float((d * v).sum())) / abs(float((g64 * v).sum()))
    idx = np.unravel_index(int(d.abs().argmax()), tuple(d.shape))
    say(f"  |g32 - g64| / |g64|: max-norm {rel_max:.3g} at {idx}, L2 {rel_l2:.3g}, "
        f"along a random direction {rel_dir:.3g}")
    # The FP32 rounding of the second-derivative stencils is of order eps |f| / dx^2, 576 times
    # its size at the 16^3 check of test_gpu_adjoint (max-norm 1e-4 there): the pointwise bound
    # is 2e-3 here (measured 5.8e-4). The L2 and directional differences average that rounding
    # out (measured 2.5e-5 and 2.2e-5); they are held to the 1e-4 of the FP32 gradient in the
    # differentiation form.
    assert rel_max < 2e-3
    assert rel_l2 < 1e-4
    assert rel_dir < 1e-4
