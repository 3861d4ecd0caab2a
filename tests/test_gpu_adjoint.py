"""GPU checks of the tangent-linear and adjoint substeps and of the gradients they produce.

  * dot-product test of one substep: <T(dx), y> = <dx, A(y)> for random directions, each s;
  * gradients of L = 0.5 |u(T) - u_target|^2 (interior) after several fixed-dt steps against
    centred finite differences in FP64 on random directions and nu, eta;
  * the same directional derivatives from the long-double host oracle at 32^3;
  * every checkpoint schedule gives the same gradient; FP32 gradients are close to FP64 ones.
"""
import ctypes
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from astaroth_amd import astaroth as ac
from astaroth_amd import adjoint as AD
import oracle.oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def physics_info(lib, n, **over):
    p = dict(AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0, AC_nu_visc=5e-3,
             AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3, AC_lupwd=True, AC_lentropy=True)
    p.update(over)
    return lib.config(n if isinstance(n, tuple) else (n, n, n), **p)


def initial(lib, info, seed):
    mesh = ac.HostMesh(lib, info)
    assert lib.acHostMeshInitFourier(ctypes.byref(mesh.struct), seed, 16, 4, 0.05, 0.0) == 0
    d = mesh.data.astype(np.float64)
    rng = np.random.default_rng(seed)
    # nonzero lnrho and ss so every term of the linearisation is exercised
    d[0] += 0.05 * np.sin(d[1] * 20) + 0.01 * rng.standard_normal(d[0].shape)
    d[7] += 0.02 * np.cos(d[5] * 30) + 0.01 * rng.standard_normal(d[7].shape)
    O.periodic_fill(d, O.config_from_info(info))
    return d.astype(lib.np_real)


def interior(t):
    return t[..., 3:-3, 3:-3, 3:-3]


def idot(a, b):
    return float((interior(a).double() * interior(b).double()).sum())


@pytest.fixture(scope="module", params=["fp64", "fp32"])
def lib(request, require_gpu):
    return ac.load(request.param)


def substep_dot_product(lib, info, f, step, entropy, seed):
    """<T(dx), y> and <dx, A(y)> of one substep at state f for seeded random directions (and
    parameter directions (0.3, -0.2)); returns (lhs, rhs, relative mismatch, param grads)."""
    # the identity is linear algebra and holds for any dt; a large one makes the Jacobian part
    # dominate both sides
    dt = lib.np_real(0.5)
    dev = ac.Device(lib, info)
    tdt = torch.float32 if lib.precision == "fp32" else torch.float64
    cuda = torch.device("cuda", 0)
    try:
        ft = torch.from_numpy(f).to(cuda)
        torch.cuda.synchronize()
        ac.check(lib.acDeviceLoadMeshFromDevicePtr(dev.handle, 0, ft.data_ptr()), "load")
        dev.periodic()
        dev.sync()
        g = torch.Generator(device="cpu").manual_seed(seed)
        rnd = lambda: torch.randn(f.shape, generator=g, dtype=torch.float64).to(tdt).to(cuda)
        df, dw, lf, lw = rnd(), rnd(), rnd(), rnd()
        if not entropy:
            df[7].zero_()
            dw[7].zero_()
            lw[7].zero_()
        if step == 2:
            lw.zero_()  # w' is dead after the last substep
        dparams = (ctypes.c_double * 2)(0.3, -0.2)
        df_out = torch.zeros_like(df)
        dw_t = dw.clone()
        torch.cuda.synchronize()
        ac.check(lib.acDeviceIntegrateSubstepTangent(dev.handle, 0, step, dt, df.data_ptr(),
                                                     df_out.data_ptr(), dw_t.data_ptr(), dparams),
                 "tangent")
        dev.sync()
        # the workspace starts as NaN bit patterns: a read of any part the adjoint did not write
        # first turns the result into NaN
        ws = torch.full((int(lib.acAdjointWorkspaceBytes(dev.handle)),), 0xFF, dtype=torch.uint8,
                        device=cuda)
        pg = torch.zeros(2, dtype=torch.float64, device=cuda)
        lf_a, lw_a = lf.clone(), lw.clone()
        torch.cuda.synchronize()
        ac.check(lib.acDeviceIntegrateSubstepAdjoint(dev.handle, 0, step, dt, lf_a.data_ptr(),
                                                     lw_a.data_ptr(), ws.data_ptr(), pg.data_ptr()),
                 "adjoint")
        dev.sync()
        lhs = idot(df_out, lf) + (idot(dw_t, lw) if step < 2 else 0.0)
        rhs = idot(df, lf_a) + (idot(dw, lw_a) if step > 0 else 0.0) + \
            0.3 * float(pg[0]) - 0.2 * float(pg[1])
        scale = abs(idot(df, lf)) + abs(lhs) + 1e-300
        return lhs, rhs, abs(lhs - rhs) / scale, pg.cpu()
    finally:
        dev.close()


@pytest.mark.parametrize("step", [0, 1, 2])
@pytest.mark.parametrize("upwd,entropy", [(True, True), (False, False), (True, False)])
# one-cell-per-thread and tiled local stage; the tiled grids' ny select the transposes' column
# block heights: 8, 4 and (FP32) 16
@pytest.mark.parametrize("n", [16, (64, 8, 16), (64, 12, 16), (64, 16, 16)])
def test_substep_dot_product(lib, step, upwd, entropy, n):
    say(f"case dot-product substep {step} {lib.precision} upwd={upwd} entropy={entropy} n={n}")
    info = physics_info(lib, n, AC_lupwd=upwd, AC_lentropy=entropy)
    f = initial(lib, info, 3 + step)
    lhs, rhs, rel, pg = substep_dot_product(lib, info, f, step, entropy, 7 + step)
    say(f"  <T dx, y> = {lhs:.15g}  <dx, A y> = {rhs:.15g}  rel = {rel:.3g}")
    tol = 1e-12 if lib.precision == "fp64" else 1e-5
    assert rel < tol
    # the parameter sensitivities alone: dot-product with dx = 0
    assert float(pg.abs().sum()) > 0


@pytest.mark.parametrize("upwd,entropy", [(True, True), (False, False), (True, False)])
@pytest.mark.parametrize("n", [16, (64, 8, 16), (64, 12, 16), (64, 16, 16)])
def test_adjoint_substep_deterministic(lib, upwd, entropy, n):
    """The adjoint substep on identical inputs gives identical outputs, bit for bit, every time.
    Regression test: the LDS-staged transposes once read a value of one plane buffer after the
    step barrier, when another wave's DMA of the next plane could already have refilled it (the
    optimiser had moved the read; the kernels now fence LDS accesses at the barrier), and the
    result then varied between runs."""
    say(f"case adjoint determinism {lib.precision} upwd={upwd} entropy={entropy} n={n}")
    info = physics_info(lib, n, AC_lupwd=upwd, AC_lentropy=entropy)
    cuda = torch.device("cuda", 0)
    tdt = torch.float32 if lib.precision == "fp32" else torch.float64
    for step in range(3):
        f = initial(lib, info, 3 + step)
        dev = ac.Device(lib, info)
        try:
            ft = torch.from_numpy(f).to(cuda)
            torch.cuda.synchronize()
            ac.check(lib.acDeviceLoadMeshFromDevicePtr(dev.handle, 0, ft.data_ptr()), "load")
            dev.periodic()
            dev.sync()
            g = torch.Generator(device="cpu").manual_seed(11 + step)
            lf = torch.randn(f.shape, generator=g, dtype=torch.float64).to(tdt).to(cuda)
            lw = torch.randn(f.shape, generator=g, dtype=torch.float64).to(tdt).to(cuda)
            ws = torch.full((int(lib.acAdjointWorkspaceBytes(dev.handle)),), 0xFF,
                            dtype=torch.uint8, device=cuda)
            outs = []
            for _ in range(6):
                lf_a, lw_a = lf.clone(), lw.clone()
                pg = torch.zeros(2, dtype=torch.float64, device=cuda)
                torch.cuda.synchronize()
                ac.check(lib.acDeviceIntegrateSubstepAdjoint(dev.handle, 0, step, lib.np_real(0.5),
                                                             lf_a.data_ptr(), lw_a.data_ptr(),
                                                             ws.data_ptr(), pg.data_ptr()),
                         "adjoint")
                dev.sync()
                outs.append((interior(lf_a).cpu(), interior(lw_a).cpu(), pg.cpu()))
            for o in outs[1:]:
                for q in range(ac.NUM_VTXBUF_HANDLES):
                    assert torch.equal(o[0][q], outs[0][0][q]), ("lambda_f", step, q)
                    assert torch.equal(o[1][q], outs[0][1][q]), ("lambda_w", step, q)
                assert torch.equal(o[2], outs[0][2]), ("params", step)
            assert torch.isfinite(outs[0][0]).all() and torch.isfinite(outs[0][1]).all()
        finally:
            dev.close()


def loss_and_target(lib, n, upwd=True):
    info = physics_info(lib, n, AC_lupwd=upwd)
    f0 = initial(lib, info, 11)
    tgt = initial(lib, info, 12)[1:4]
    return info, f0, tgt


def run_loss(runner, f0, tgt, nu, eta):
    out, _ = runner.run_forward(torch.from_numpy(f0).to(runner.tdev), nu, eta)
    u = interior(out[1:4]).double()
    t = interior(torch.from_numpy(tgt).to(out.device)).double()
    return 0.5 * float(((u - t) ** 2).sum())


def adjoint_gradient(runner, f0, tgt, nu0, eta0):
    tdt = runner.tdtype
    f0t = torch.from_numpy(f0).to(runner.tdev).requires_grad_(True)
    nu = torch.tensor(nu0, dtype=tdt, requires_grad=True)
    eta = torch.tensor(eta0, dtype=tdt, requires_grad=True)
    out = AD.MHDSolve.apply(f0t, nu, eta, runner)
    t = torch.from_numpy(tgt).to(runner.tdev)
    L = 0.5 * ((interior(out[1:4]) - interior(t)) ** 2).sum()
    L.backward()
    return f0t.grad.detach(), float(nu.grad), float(eta.grad)


def random_direction(f0, seed):
    rng = np.random.default_rng(seed)
    v = np.zeros_like(f0)
    v[:, 3:-3, 3:-3, 3:-3] = 0.01 * rng.standard_normal(interior(v).shape)
    return v


# Step of the centred differences. The loss is smooth without upwinding, and h = 1e-3 resolves
# its derivative to ~1e-8. The upwind term |u_j| up_j(f) has a kink wherever a velocity
# component crosses zero; differences over h = 1e-3 straddle such cells and miss the one-sided
# derivative by ~2e-6 (relative), decreasing linearly with h (measured with the long-double
# oracle: 1e-3 -> 2.2e-6, 3e-4 -> 6e-7, 1e-4 -> 1e-9), so with upwinding h = 1e-5 is used.
FD_STEP = {False: 1e-3, True: 1e-5}


@pytest.mark.parametrize("upwd", [False, True])
@pytest.mark.parametrize("schedule", ["revolve", "uniform"])
def test_gradient_vs_finite_differences_fp64(schedule, upwd, require_gpu):
    lib = ac.load("fp64")
    n, nsteps, dt = 16, 6, 2e-2
    say(f"case gradient vs FD fp64 schedule={schedule} upwd={upwd}")
    info, f0, tgt = loss_and_target(lib, n, upwd)
    runner = AD.MHDAdjoint(lib, info, nsteps, dt, schedule=schedule, snaps=2, every=4)
    try:
        nu0, eta0 = 5e-3, 5e-3
        g, gnu, geta = adjoint_gradient(runner, f0, tgt, nu0, eta0)
        assert float(g[:, :3].abs().sum()) == 0.0  # halos get no gradient
        g = g.cpu().numpy()
        h = FD_STEP[upwd]
        for k in range(3):
            v = random_direction(f0, 5 + k)
            fd = (run_loss(runner, f0 + h * v, tgt, nu0, eta0) -
                  run_loss(runner, f0 - h * v, tgt, nu0, eta0)) / (2 * h)
            ad = float((g * v).sum())
            rel = abs(fd - ad) / abs(fd)
            say(f"  direction {k}: FD {fd:.12g} adjoint {ad:.12g} rel {rel:.3g}")
            assert rel < 1e-6
        for name, grad in (("nu", gnu), ("eta", geta)):
            hp = 1e-3 * (nu0 if name == "nu" else eta0)
            args = lambda sgn: (nu0 + sgn * hp, eta0) if name == "nu" else (nu0, eta0 + sgn * hp)
            fd = (run_loss(runner, f0, tgt, *args(1)) - run_loss(runner, f0, tgt, *args(-1))) / (2 * hp)
            rel = abs(fd - grad) / abs(fd)
            say(f"  d/d{name}: FD {fd:.12g} adjoint {grad:.12g} rel {rel:.3g}")
            assert rel < 1e-6
        say(f"  forward steps {runner.forward_steps}, adjoint steps {runner.adjoint_steps}")
    finally:
        runner.close()


def test_schedules_agree(require_gpu):
    """revolve with few snapshots, revolve with many, and uniform checkpoints give the same
    gradient bit for bit (the schedules only change which states are recomputed). The solve
    kernel is pinned so that every runner's device computes the forward states alike."""
    lib = ac.load("fp64")
    say("case checkpoint schedules agree")
    info, f0, tgt = loss_and_target(lib, 16)
    info["AC_solve_tile_x"] = -1
    grads = []
    for sched, snaps, every in (("revolve", 1, 0), ("revolve", 9, 0), ("uniform", 0, 3)):
        runner = AD.MHDAdjoint(lib, info, 9, 2e-2, schedule=sched, snaps=snaps, every=every or 10)
        try:
            grads.append(adjoint_gradient(runner, f0, tgt, 5e-3, 5e-3))
            say(f"  {sched} snaps={snaps} every={every}: forward steps {runner.forward_steps}")
        finally:
            runner.close()
    for g in grads[1:]:
        assert torch.equal(g[0], grads[0][0])
        assert g[1] == grads[0][1] and g[2] == grads[0][2]


@pytest.mark.parametrize("upwd", [False, True])
def test_gradient_vs_host_oracle_fp64(upwd, require_gpu):
    """Directional derivatives of the loss at 32^3 from the long-double host oracle (centred
    differences, FP64 state storage) against the GPU adjoint gradient."""
    lib = ac.load("fp64")
    n, nsteps, dt = 32, 3, 1e-2
    say(f"case gradient vs host oracle fp64 32^3 upwd={upwd}")
    info, f0, tgt = loss_and_target(lib, n, upwd)
    cfg = O.config_from_info(info)
    runner = AD.MHDAdjoint(lib, info, nsteps, dt, schedule="revolve", snaps=1)
    try:
        g, _, _ = adjoint_gradient(runner, f0, tgt, 5e-3, 5e-3)
        g = g.cpu().numpy()
    finally:
        runner.close()

    def host_loss(f):
        m = f.copy()
        for _ in range(nsteps):
            O.integrate_step(m, cfg, dt)
        return 0.5 * float(((interior(m[1:4]) - interior(tgt)) ** 2).sum())

    h = FD_STEP[upwd]
    v = random_direction(f0, 9)
    fd = (host_loss(f0 + h * v) - host_loss(f0 - h * v)) / (2 * h)
    ad = float((g * v).sum())
    rel = abs(fd - ad) / abs(fd)
    say(f"  host oracle FD {fd:.12g} adjoint {ad:.12g} rel {rel:.3g}")
    assert rel < 1e-6


def test_gradient_fp32_close_to_fp64(require_gpu):
    n, nsteps, dt = 16, 4, 2e-2
    say("case fp32 gradient vs fp64 gradient")
    grads = {}
    for prec in ("fp64", "fp32"):
        lib = ac.load(prec)
        info, f0, tgt = loss_and_target(lib, n)
        runner = AD.MHDAdjoint(lib, info, nsteps, dt, schedule="revolve", snaps=2)
        try:
            g, gnu, geta = adjoint_gradient(runner, f0, tgt, 5e-3, 5e-3)
            grads[prec] = (g.double().cpu().numpy(), gnu, geta)
        finally:
            runner.close()
    g64, g32 = grads["fp64"][0], grads["fp32"][0]
    rel = np.abs(g64 - g32).max() / np.abs(g64).max()
    say(f"  max |g32 - g64| / max |g64| = {rel:.3g}")
    assert rel < 1e-4
    for k, name in ((1, "nu"), (2, "eta")):
        r = abs(grads["fp32"][k] - grads["fp64"][k]) / abs(grads["fp64"][k])
        say(f"  d/d{name}: fp32 {grads['fp32'][k]:.8g} fp64 {grads['fp64'][k]:.8g} rel {r:.3g}")
        assert r < 1e-3
