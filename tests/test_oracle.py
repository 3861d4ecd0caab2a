"""Pins of the CPU oracle (oracle/oracle.cc) against closed forms and independent statements.

Each check is chosen so that a plausible slip in the oracle (a coefficient, a sign, an index,
a transposed operand, a dropped term, a wrong RK coefficient) fails it:
  * stencil symbols of the first, second, mixed and upwind operators, isolated through terms of
    the right-hand side that are linear in one field;
  * the full nonlinear right-hand side against tests/physics_np.py on single-mode fields;
  * RK3 linear amplification 1 + z + z^2/2 + z^3/6 (shear wave, Beltrami A-field);
  * invariance of a uniform state; periodic fill by brute force; reductions on closed forms;
  * the FP64 ref-cpu baseline against the long-double path.
"""
import numpy as np
import pytest

import oracle.oracle as O
from tests import physics_np as PN

N = 16
H = 2 * np.pi / N


def coords(n=N, h=H):
    m = n + 6
    x = (np.arange(m) - 3) * h
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def empty(n=N):
    return np.zeros((8, n + 6, n + 6, n + 6))


def interior(a, n=N):
    return a[..., 3:n + 3, 3:n + 3, 3:n + 3]


def rhs_of(f, cfg):
    f = np.ascontiguousarray(f)
    O.periodic_fill(f, cfg)
    return O.rhs(f, cfg)


@pytest.mark.parametrize("axis,k", [(0, 1), (0, 3), (1, 2), (2, 5)])
def test_first_derivative_symbol(axis, k):
    """Isothermal pressure term: u = A = 0, lnrho = a sin(k x_axis) -> du/dt = -cs^2 grad lnrho."""
    cfg = O.make_config((N, N, N), H, lentropy=0, cs_sound=1.3)
    X = coords()
    f = empty()
    a = 0.7
    f[0] = a * np.sin(k * X[axis])
    r = rhs_of(f, cfg)
    expect = -1.3 ** 2 * a * PN.sym_first(k, H) * np.cos(k * X[axis])
    np.testing.assert_allclose(interior(r[1 + axis]), interior(expect), rtol=0, atol=1e-12)
    for q in range(8):
        if q != 1 + axis:
            np.testing.assert_allclose(interior(r[q]), 0.0, atol=1e-12)
    # a sign or coefficient slip in D1 would move this by >> tolerance
    assert abs(PN.sym_first(k, H) - k) < 0.5 * k


@pytest.mark.parametrize("k", [1, 2, 4])
def test_second_derivative_symbol(k):
    """Shear wave u_x = a sin(k y): du_x/dt = nu d_yy u_x (all other terms vanish)."""
    nu = 0.05
    cfg = O.make_config((N, N, N), H, lentropy=0, nu_visc=nu)
    X = coords()
    f = empty()
    f[1] = 0.4 * np.sin(k * X[1])
    r = rhs_of(f, cfg)
    expect = nu * PN.sym_second(k, H) * f[1]
    np.testing.assert_allclose(interior(r[1]), interior(expect), atol=1e-12)
    np.testing.assert_allclose(interior(r[2]), 0.0, atol=1e-12)


@pytest.mark.parametrize("pair,ks", [((0, 1), (1, 2)), ((0, 2), (3, 1)), ((1, 2), (2, 2))])
def test_mixed_derivative_symbol(pair, ks):
    """u_a = sin(ka x_a + kb x_b): du_b/dt = (nu/3 + zeta) d_a d_b u_a (grad-div term)."""
    nu, zeta = 0.02, 0.01
    cfg = O.make_config((N, N, N), H, lentropy=0, nu_visc=nu, zeta=zeta)
    X = coords()
    f = empty()
    a, b = pair
    f[1 + a] = 0.3 * np.sin(ks[0] * X[a] + ks[1] * X[b])
    r = rhs_of(f, cfg)
    expect = (nu / 3 + zeta) * PN.sym_mixed(ks[0], ks[1], H, H) * f[1 + a]
    np.testing.assert_allclose(interior(r[1 + b]), interior(expect), atol=1e-12)


@pytest.mark.parametrize("U0", [0.8, -0.5])
def test_upwind_symbol(U0):
    """Uniform flow u_x = U0 over lnrho = a sin(kx): dlnrho/dt = -U0 d_x lnrho + |U0| upw_x."""
    k = 3
    cfg = O.make_config((N, N, N), H, lentropy=0, lupwd=1)
    X = coords()
    f = empty()
    f[0] = 0.2 * np.sin(k * X[0])
    f[1] = U0
    r = rhs_of(f, cfg)
    expect = (-U0 * 0.2 * PN.sym_first(k, H) * np.cos(k * X[0]) +
              abs(U0) * PN.sym_upwind(k, H) * f[0])
    np.testing.assert_allclose(interior(r[0]), interior(expect), atol=1e-12)
    cfg0 = O.make_config((N, N, N), H, lentropy=0, lupwd=0)
    r0 = rhs_of(f, cfg0)
    np.testing.assert_allclose(interior(r0[0]), interior(-U0 * 0.2 * PN.sym_first(k, H) *
                                                         np.cos(k * X[0])), atol=1e-12)


@pytest.mark.parametrize("entropy,upwd", [(1, 1), (0, 1), (1, 0)])
def test_full_rhs_single_modes(entropy, upwd):
    """Whole nonlinear right-hand side vs the independent numpy statement."""
    rng = np.random.default_rng(7 + entropy + 2 * upwd)
    P = dict(cs_sound=1.1, nu=0.02, zeta=0.013, eta=0.017, mu0=1.4, cp=1.2, gamma=5 / 3,
             lnT0=1.2, lnrho0=1.3, K=0.011)
    cfg = O.make_config((N, N, N), H, cs_sound=P["cs_sound"], nu_visc=P["nu"], zeta=P["zeta"],
                        eta=P["eta"], mu0=P["mu0"], cp_sound=P["cp"], gamma=P["gamma"],
                        lnT0=P["lnT0"], lnrho0=P["lnrho0"], K_heatcond=P["K"], lupwd=upwd,
                        lentropy=entropy)
    X = coords()
    modes = []
    f = empty()
    for q in range(8):
        k = rng.integers(-3, 4, size=3)
        if not k.any():
            k[0] = 1
        a = 0.3 * rng.random() + 0.05
        phi = 2 * np.pi * rng.random()
        modes.append((a, k, phi))
        f[q] = a * np.sin(k[0] * X[0] + k[1] * X[1] + k[2] * X[2] + phi)
    r = rhs_of(f, cfg)
    expect = PN.rhs_modes(modes, X, (H, H, H), P, upwd=bool(upwd), entropy=bool(entropy))
    for q in range(8):
        e = interior(expect[q])
        scale = max(np.max(np.abs(e)), 1e-30)
        np.testing.assert_allclose(interior(r[q]) / scale, e / scale, atol=2e-12,
                                   err_msg=f"field {q}")


def p3(z):
    return 1 + z + z * z / 2 + z ** 3 / 6


@pytest.mark.parametrize("dtype", [np.float64])
def test_rk3_amplification_shear_wave(dtype):
    nu, k, dt, steps = 0.05, 3, 0.05, 3
    cfg = O.make_config((N, N, N), H, lentropy=0, nu_visc=nu)
    X = coords()
    f = empty().astype(dtype)
    f[1] = 0.4 * np.sin(k * X[1])
    f0 = f[1].copy()
    for _ in range(steps):
        O.integrate_step(f, cfg, dt)
    z = nu * PN.sym_second(k, H) * dt
    assert abs(z ** 3 / 6) > 1e-6  # the cubic term matters at this z
    np.testing.assert_allclose(interior(f[1]), interior(p3(z) ** steps * f0), atol=1e-13)
    # a second-order polynomial would be off by far more than the tolerance
    assert np.max(np.abs(interior((1 + z + z * z / 2) ** steps * f0 - p3(z) ** steps * f0))) > 1e-7


def test_beltrami_decay_with_entropy():
    """A = a(0, sin kx, cos kx): j || B so j x B = 0; A decays as P3(eta k2 dt)^n; u stays 0."""
    eta, k, dt, steps = 0.04, 2, 0.05, 2
    cfg = O.make_config((N, N, N), H, lentropy=1, eta=eta, gamma=5 / 3)
    X = coords()
    f = empty()
    a = 0.25
    f[5] = a * np.sin(k * X[0])
    f[6] = a * np.cos(k * X[0])
    f0 = f.copy()
    for _ in range(steps):
        O.integrate_step(f, cfg, dt)
    z = eta * PN.sym_second(k, H) * dt
    for q in (5, 6):
        np.testing.assert_allclose(interior(f[q]), interior(p3(z) ** steps * f0[q]), atol=1e-13)
    for q in (1, 2, 3, 4):
        assert np.max(np.abs(interior(f[q]))) < 1e-14
    # Ohmic heating is uniform: ss rises by the same amount everywhere
    s = interior(f[7])
    assert np.max(s) > 0 and np.ptp(s) < 1e-12 * np.max(s)


def test_uniform_state_invariant():
    cfg = O.make_config((N, N, N), H)
    f = empty()
    vals = [0.3, 0.1, -0.2, 0.05, 0.7, -0.4, 0.2, 0.15]
    for q in range(8):
        f[q] = vals[q]
    f0 = f.copy()
    O.integrate_step(f, cfg, 1e-2)
    np.testing.assert_allclose(interior(f), interior(f0), atol=1e-14)


def test_periodic_fill_bruteforce():
    n = (5, 7, 6)
    cfg = O.make_config(n, 0.1)
    rng = np.random.default_rng(3)
    f = rng.random((8, n[2] + 6, n[1] + 6, n[0] + 6))
    O.periodic_fill(f, cfg)
    zz = (np.arange(n[2] + 6) - 3) % n[2] + 3
    yy = (np.arange(n[1] + 6) - 3) % n[1] + 3
    xx = (np.arange(n[0] + 6) - 3) % n[0] + 3
    expect = f[:, zz][:, :, yy][:, :, :, xx]
    np.testing.assert_array_equal(f, expect)


def test_reductions_closed_form():
    cfg = O.make_config((N, N, N), H)
    X = coords()
    f = np.ascontiguousarray(np.sin(2 * X[0]))
    # RMS of sin over a periodic grid is exactly 1/sqrt(2)
    rms = float(O.reduce([f], cfg, 2, 3, True))
    assert abs(rms - np.sqrt(0.5)) < 1e-15
    assert abs(float(O.reduce([f], cfg, 1, 2, False)) - np.max(interior(f))) == 0
    assert abs(float(O.reduce([f], cfg, 1, 1, False)) - np.min(interior(f))) == 0
    g = np.ascontiguousarray(np.cos(2 * X[0]))
    nrm = float(O.reduce([f, g, np.zeros_like(f)], cfg, 4, 2, False))
    assert abs(nrm - 1.0) < 1e-15
    lnr = np.ascontiguousarray(np.log(4.0) * np.ones_like(f))
    alf = float(O.reduce([f, g, np.zeros_like(f), lnr], cfg, 7, 2, False))
    assert abs(alf - 0.5) < 1e-15


def test_refcpu_matches_long_double_path():
    rng = np.random.default_rng(11)
    cfg = O.make_config((N, N, N), H, nu_visc=5e-3, eta=5e-3, K_heatcond=5e-3, gamma=5 / 3)
    X = coords()
    f = empty()
    for q in range(1, 7):
        k = rng.integers(-2, 3, size=3)
        f[q] = 0.05 * np.sin(k[0] * X[0] + k[1] * X[1] + k[2] * X[2] + rng.random() * 6)
    g = f.copy()
    w = np.zeros_like(g)
    for _ in range(2):
        O.integrate_step(f, cfg, 1e-2)
        O.refcpu_step(g, w, cfg, 1e-2)
    for q in range(8):
        scale = max(np.max(np.abs(interior(f[q]))), 1e-30)
        assert np.max(np.abs(interior(f[q] - g[q]))) / scale < 1e-12, q


def test_refcpu_fp32_tracks_long_double_path():
    """The reference CPU build's algorithm in FP32 arithmetic and storage (AcReal = float) stays
    within FP32 rounding of the long-double path (with float storage), and differs from it by
    more than a double computation would."""
    rng = np.random.default_rng(12)
    cfg = O.make_config((N, N, N), H, nu_visc=5e-3, eta=5e-3, K_heatcond=5e-3, gamma=5 / 3)
    X = coords()
    f = empty()
    for q in range(1, 7):
        k = rng.integers(-2, 3, size=3)
        f[q] = 0.05 * np.sin(k[0] * X[0] + k[1] * X[1] + k[2] * X[2] + rng.random() * 6)
    f32 = f.astype(np.float32)
    model = f32.copy()
    g = f32.copy()
    w = np.zeros_like(g)
    for _ in range(2):
        O.integrate_step(model, cfg, 1e-2)
        O.refcpu_step(g, w, cfg, 1e-2)
    worst = 0.0
    for q in range(1, 7):
        scale = np.max(np.abs(interior(model[q]).astype(np.float64)))
        err = np.max(np.abs(interior(model[q]).astype(np.float64) - interior(g[q]))) / scale
        worst = max(worst, err)
        assert err < 1e-5, (q, err)
    # float arithmetic leaves errors well above the storage rounding of a double computation
    assert worst > 1e-8, worst


def test_float_storage_close_to_double():
    rng = np.random.default_rng(5)
    cfg = O.make_config((N, N, N), H)
    f = empty()
    X = coords()
    for q in range(1, 7):
        f[q] = 0.05 * np.cos(X[0] * (q % 3 + 1) + X[1] * (q % 2) + rng.random())
    g = f.astype(np.float32)
    O.integrate_step(f, cfg, 1e-3)
    O.integrate_step(g, cfg, 1e-3)
    assert np.max(np.abs(interior(f) - interior(g))) < 1e-7


def test_error_metric():
    e = O.get_error(1.0, 1.0 + 2 ** -23, 24)
    assert abs(e["ulp"] - 1.0) < 1e-12
    e = O.get_error(3.0, 3.0 + 2 * 2 ** -22, 24)
    assert abs(e["ulp"] - 2.0) < 1e-12
    assert O.get_error(2.0, 2.0, 53)["ulp"] == 0
