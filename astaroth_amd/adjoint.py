"""Gradients of the MHD solve through the hand-written adjoint substep (no reference counterpart).

`MHDAdjoint` runs `nsteps` RK3 steps of fixed dt forward on the device and back-propagates an
adjoint state through them with `acDeviceIntegrateSubstepAdjoint`, recomputing forward states
from checkpoints. Checkpoint schedules:

  * "revolve": binomial checkpointing with `snaps` stored states (Griewank & Walther's revolve
    rule, here found by dynamic programming over the split point of each segment, which yields
    the same minimal number of recomputed steps);
  * "uniform": a checkpoint every `every` steps; each segment is recomputed once with all of its
    steps stored (two-level checkpointing).

`MHDSolve.apply(f0, nu, eta, runner)` is the `torch.autograd.Function`: f0 is the full state
(NUM_VTXBUF_HANDLES, mz, my, mx) on the GPU, nu and eta are 0-d tensors; it returns the state after
the run with periodic halos filled, and its backward gives gradients for f0 (interior cells;
halos receive none, since the first halo fill overwrites them), nu and eta.

torch provides device memory and the synchronisation points here; every computation runs in the
HIP kernels of astaroth_amd/lib through the C ABI.
"""
import ctypes
import functools

import numpy as np
import torch

from . import astaroth as ac

NF = ac.NUM_VTXBUF_HANDLES


@functools.lru_cache(maxsize=None)
def _revolve_table(steps, snaps):
    """cost[l][c]: fewest forward steps to reverse l steps with c free snapshot slots (the state
    at the start of the segment is held separately); split[l][c]: the first advance."""
    INF = 1 << 60
    cost = np.full((steps + 1, snaps + 1), INF, dtype=np.int64)
    split = np.zeros((steps + 1, snaps + 1), dtype=np.int64)
    for c in range(snaps + 1):
        cost[1][c] = 0
    for l in range(2, steps + 1):
        cost[l][0] = l * (l - 1) // 2  # recompute from the start before every reverse step
        split[l][0] = l - 1
        for c in range(1, snaps + 1):
            t = np.arange(1, l)
            tot = t + cost[l - t, c - 1] + cost[t, c]
            k = int(np.argmin(tot))
            cost[l][c] = tot[k]
            split[l][c] = t[k]
    return cost, split


def revolve_actions(steps, snaps):
    """Action list for reversing `steps` steps with `snaps` snapshot slots besides the initial
    state: ("advance", i, j) runs steps i..j-1 forward from the state at i (held by the caller),
    ("store", j, slot) / ("restore", i, slot) / ("free", slot) move states, ("reverse", i) runs
    the adjoint of step i (its input state is current)."""
    _, split = _revolve_table(steps, snaps)
    acts = []

    def rec(i0, i1, slot0, free):
        l = i1 - i0
        if l == 1:
            acts.append(("restore", i0, slot0))
            acts.append(("reverse", i0))
            return
        c = len(free)
        t = int(split[l][c])
        if c == 0:
            # no slot: recompute from i0 before each reverse step
            for i in range(i1 - 1, i0 - 1, -1):
                acts.append(("restore", i0, slot0))
                if i > i0:
                    acts.append(("advance", i0, i))
                acts.append(("reverse", i))
            return
        m = i0 + t
        slot = free[0]
        acts.append(("restore", i0, slot0))
        acts.append(("advance", i0, m))
        acts.append(("store", m, slot))
        rec(m, i1, slot, free[1:])
        acts.append(("free", slot))
        rec(i0, m, slot0, free)

    rec(0, steps, "init", list(range(snaps)))
    return acts


def uniform_actions(steps, every):
    """Checkpoints every `every` steps (taken during the forward run); each segment is
    recomputed once with all of its step states stored, then reversed."""
    acts = []
    starts = list(range(0, steps, every))
    for s0 in reversed(starts):
        s1 = min(s0 + every, steps)
        acts.append(("restore", s0, ("ckpt", s0)))
        for i in range(s0, s1):
            if i > s0:
                acts.append(("advance", i - 1, i))
            acts.append(("store", i, ("seg", i)))
        for i in range(s1 - 1, s0 - 1, -1):
            acts.append(("restore", i, ("seg", i)))
            acts.append(("reverse", i))
    return acts


class MHDAdjoint:
    """Forward/adjoint runner on one device for a fixed-dt integration of `nsteps` steps."""

    def __init__(self, lib, info, nsteps, dt, schedule="revolve", snaps=8, every=10,
                 device_id=0):
        if info["AC_kahan_bits"]:
            raise ValueError("the differentiable path runs without Kahan compensation")
        self.lib = lib
        self.dev = ac.Device(lib, info, device_id)
        self.info = self.dev.info
        self.nsteps = nsteps
        self.dt = lib.np_real(dt)
        self.schedule = schedule
        self.snaps = snaps
        self.every = every
        mx, my, mz = self.info["AC_mlocal"]
        self.shape = (NF, mz, my, mx)
        self.count = mx * my * mz
        self.tdtype = torch.float32 if lib.precision == "fp32" else torch.float64
        self.tdev = torch.device("cuda", device_id)
        ws = lib.acAdjointWorkspaceBytes(self.dev.handle)
        self.workspace = torch.empty(int(ws), dtype=torch.uint8, device=self.tdev)
        self.param_grad = torch.zeros(2, dtype=torch.float64, device=self.tdev)
        self.forward_steps = 0   # forward steps run (including recomputation), for cost reports
        self.adjoint_steps = 0

    # ---- state movement between torch tensors and the device
    def _sync(self):
        ac.check(self.lib.acDeviceSynchronizeStream(self.dev.handle, ac.STREAM_ALL), "sync")

    def load(self, t):
        torch.cuda.synchronize(self.tdev)
        ac.check(self.lib.acDeviceLoadMeshFromDevicePtr(self.dev.handle, 0, t.data_ptr()), "load")
        self._sync()

    def store(self, t):
        self._sync()
        ac.check(self.lib.acDeviceStoreMeshToDevicePtr(self.dev.handle, 0, t.data_ptr()), "store")
        self._sync()

    def empty(self):
        return torch.empty(self.shape, dtype=self.tdtype, device=self.tdev)

    def set_params(self, nu, eta):
        ac.check(self.lib.acDeviceLoadScalarUniform(self.dev.handle, 0, ac.AC_nu_visc, float(nu)), "nu")
        ac.check(self.lib.acDeviceLoadScalarUniform(self.dev.handle, 0, ac.AC_eta, float(eta)), "eta")

    def advance(self, k):
        for _ in range(k):
            self.dev.step(self.dt)
        self.forward_steps += k

    # ---- forward
    def run_forward(self, f0, nu, eta):
        """Runs nsteps from f0; returns the final state (halos filled) and the checkpoints."""
        self.set_params(nu, eta)
        self.load(f0.contiguous())
        ckpts = {}
        if self.schedule == "uniform":
            for i in range(self.nsteps):
                if i % self.every == 0:
                    t = self.empty()
                    self.store(t)
                    ckpts[("ckpt", i)] = t
                self.advance(1)
        else:
            ckpts["init"] = f0.detach().clone()
            self.advance(self.nsteps)
        self.dev.periodic()
        out = self.empty()
        self.store(out)
        return out, ckpts

    # ---- adjoint of one step whose input state is in the device
    def reverse_step(self, lam_f, lam_w):
        # the states entering substeps 0 and 1 are saved; the one entering substep 2 is left in
        # the device by the recomputation, where the first adjoint substep reads it (an adjoint
        # substep reads the device state and writes only lambda and its workspace)
        s0 = self.empty()
        self.store(s0)
        states = [s0]
        for s in range(2):
            self.dev.periodic()
            self.dev.substep(s, self.dt)
            self.dev.swap()
            if s == 0:
                t = self.empty()
                self.store(t)
                states.append(t)
        for s in (2, 1, 0):
            if s < 2:
                self.load(states[s])
            self.dev.periodic()
            ac.check(self.lib.acDeviceIntegrateSubstepAdjoint(
                self.dev.handle, 0, s, self.dt, ctypes.c_void_p(lam_f.data_ptr()),
                ctypes.c_void_p(lam_w.data_ptr()), ctypes.c_void_p(self.workspace.data_ptr()),
                ctypes.c_void_p(self.param_grad.data_ptr())), "adjoint substep")
        self._sync()
        self.adjoint_steps += 1

    def run_adjoint(self, lam_T, ckpts, nu, eta):
        """lam_T: dL/d(final state) on interior cells. Returns (dL/df0, dL/dnu, dL/deta)."""
        self.set_params(nu, eta)
        lam_f = fold_halo_transpose(lam_T.to(self.tdtype)).contiguous()
        lam_w = torch.zeros_like(lam_f)
        self.param_grad.zero_()
        torch.cuda.synchronize(self.tdev)
        store = dict(ckpts)
        acts = (uniform_actions(self.nsteps, self.every) if self.schedule == "uniform"
                else revolve_actions(self.nsteps, self.snaps))
        for act in acts:
            kind = act[0]
            if kind == "restore":
                self.load(store[act[2]])
            elif kind == "advance":
                self.advance(act[2] - act[1])
            elif kind == "store":
                t = self.empty()
                self.store(t)
                store[act[2]] = t
            elif kind == "free":
                store.pop(act[1], None)
            elif kind == "reverse":
                self.reverse_step(lam_f, lam_w)
        torch.cuda.synchronize(self.tdev)
        g = zero_halos(lam_f)
        pg = self.param_grad.detach().cpu().numpy()
        return g, float(pg[0]), float(pg[1])

    def close(self):
        self.dev.close()


def fold_halo_transpose(lam):
    """Transpose of the periodic halo fill: halo values are added into the interior cells they
    are copies of, then cleared (per axis; the 3D fill is the composition of the three)."""
    lam = lam.clone()
    r = 3
    for axis in (3, 2, 1):  # x, y, z of (field, z, y, x)
        m = lam.shape[axis]
        n = m - 2 * r
        lo = lam.narrow(axis, 0, r)                 # copies of interior [n, n + r)
        hi = lam.narrow(axis, n + r, r)             # copies of interior [r, 2r)
        lam.narrow(axis, n, r).add_(lo)
        lam.narrow(axis, r, r).add_(hi)
        lo.zero_()
        hi.zero_()
    return lam


def zero_halos(t):
    out = torch.zeros_like(t)
    r = 3
    out[:, r:-r, r:-r, r:-r] = t[:, r:-r, r:-r, r:-r]
    return out


class MHDSolve(torch.autograd.Function):
    """f_T = solve(f_0; nu, eta) over runner.nsteps fixed-dt RK3 steps, differentiable."""

    @staticmethod
    def forward(ctx, f0, nu, eta, runner):
        out, ckpts = runner.run_forward(f0.detach(), float(nu), float(eta))
        ctx.runner = runner
        ctx.ckpts = ckpts
        ctx.params = (float(nu), float(eta))
        ctx.dtypes = (nu.dtype, eta.dtype)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        runner = ctx.runner
        g, gnu, geta = runner.run_adjoint(grad_out.detach(), ctx.ckpts, *ctx.params)
        dev = grad_out.device
        return (g, torch.tensor(gnu, dtype=ctx.dtypes[0], device=dev),
                torch.tensor(geta, dtype=ctx.dtypes[1], device=dev), None)
