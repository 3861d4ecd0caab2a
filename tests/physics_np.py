"""Independent numpy statement of the MHD right-hand side for single-Fourier-mode fields.

Every field q is f_q = a_q sin(k_q . x + phi_q). For such fields the 6th-order discrete operators
act exactly as multipliers (their Fourier symbols), so the discrete right-hand side has a closed
form. This module evaluates that closed form from the governing equations (continuity,
momentum, induction, entropy with heat conduction: ref acc-runtime/samples/mhd_modular/
mhdsolver.ac:279-426), written directly in numpy. It shares no code with the oracle and is used
to pin the oracle's full nonlinear right-hand side.
"""
import numpy as np

C1 = np.array([0.0, 3 / 4, -3 / 20, 1 / 60])
C2 = np.array([-49 / 18, 3 / 2, -3 / 20, 1 / 90])
CM = np.array([0.0, 270 / 720, -27 / 720, 2 / 720])


def sym_first(k, h):
    th = k * h
    return (1.5 * np.sin(th) - 0.3 * np.sin(2 * th) + np.sin(3 * th) / 30.0) / h


def sym_second(k, h):
    th = k * h
    return (-49 / 18 + 3 * np.cos(th) - 0.3 * np.cos(2 * th) + np.cos(3 * th) / 45.0) / h ** 2


def sym_mixed(ka, kb, ha, hb):
    s = 0.0
    for i in (1, 2, 3):
        s += CM[i] * np.sin(i * ka * ha) * np.sin(i * kb * hb)
    return -4.0 * s / (ha * hb)


def sym_upwind(k, h):
    return -(8.0 / (60.0 * h)) * (1 - np.cos(k * h)) ** 3


def mode_derivs(a, k, phi, X, h):
    """value, gradient[3], hessian[3][3], upwind[3] of a sin(k.x + phi) on coordinates X."""
    ph = k[0] * X[0] + k[1] * X[1] + k[2] * X[2] + phi
    s, c = a * np.sin(ph), a * np.cos(ph)
    grad = [sym_first(k[d], h[d]) * c for d in range(3)]
    hess = [[None] * 3 for _ in range(3)]
    for d in range(3):
        hess[d][d] = sym_second(k[d], h[d]) * s
    for (p, q) in ((0, 1), (0, 2), (1, 2)):
        hess[p][q] = hess[q][p] = sym_mixed(k[p], k[q], h[p], h[q]) * s
    upw = [sym_upwind(k[d], h[d]) * s for d in range(3)]
    return s, grad, hess, upw


def rhs_modes(modes, X, h, P, upwd=True, entropy=True):
    """modes: list of 8 (a, k(3,), phi). P: dict of physics parameters. Returns list of 8 arrays."""
    D = [mode_derivs(a, np.asarray(k, float), phi, X, h) for (a, k, phi) in modes]
    lr, glr, hlr, ulr = D[0]
    u = [D[1 + i][0] for i in range(3)]
    gu = [D[1 + i][1] for i in range(3)]       # gu[i][j] = d_j u_i
    hu = [D[1 + i][2] for i in range(3)]
    uu_up = [D[1 + i][3] for i in range(3)]
    ga = [D[4 + i][1] for i in range(3)]
    ha = [D[4 + i][2] for i in range(3)]
    s, gs, hs, _ = D[7]

    lap = lambda H: H[0][0] + H[1][1] + H[2][2]
    divu = gu[0][0] + gu[1][1] + gu[2][2]
    # grad(div v)_i = sum_j d_i d_j v_j
    gdu = [sum(hu[j][i][j] for j in range(3)) for i in range(3)]
    gda = [sum(ha[j][i][j] for j in range(3)) for i in range(3)]
    lapu = [lap(hu[i]) for i in range(3)]
    lapa = [lap(ha[i]) for i in range(3)]
    absu = [np.abs(x) for x in u]

    cont = -(u[0] * glr[0] + u[1] * glr[1] + u[2] * glr[2]) - divu
    if upwd:
        cont = cont + absu[0] * ulr[0] + absu[1] * ulr[1] + absu[2] * ulr[2]

    S = [[None] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            S[i][j] = 0.5 * (gu[i][j] + gu[j][i])
        S[i][i] = S[i][i] - divu / 3.0
    B = [ga[2][1] - ga[1][2], ga[0][2] - ga[2][0], ga[1][0] - ga[0][1]]
    J = [(gda[i] - lapa[i]) / P["mu0"] for i in range(3)]
    JxB = [J[1] * B[2] - J[2] * B[1], J[2] * B[0] - J[0] * B[2], J[0] * B[1] - J[1] * B[0]]
    rho = np.exp(lr)
    cs2 = P["cs_sound"] ** 2
    if entropy:
        cs2 = cs2 * np.exp(P["gamma"] * s / P["cp"] + (P["gamma"] - 1) * (lr - P["lnrho0"]))
    mom = []
    for i in range(3):
        m = -(gu[i][0] * u[0] + gu[i][1] * u[1] + gu[i][2] * u[2])
        if upwd:
            m = m + uu_up[i][0] * absu[0] + uu_up[i][1] * absu[1] + uu_up[i][2] * absu[2]
        gp = (gs[i] / P["cp"] + glr[i]) if entropy else glr[i]
        m = m - cs2 * gp + JxB[i] / rho
        m = m + P["nu"] * (lapu[i] + gdu[i] / 3.0 + 2.0 * (S[i][0] * glr[0] + S[i][1] * glr[1] +
                                                          S[i][2] * glr[2]))
        m = m + P["zeta"] * gdu[i]
        mom.append(m)
    ind = [u[1] * B[2] - u[2] * B[1] + P["eta"] * lapa[0],
           u[2] * B[0] - u[0] * B[2] + P["eta"] * lapa[1],
           u[0] * B[1] - u[1] * B[0] + P["eta"] * lapa[2]]
    if entropy:
        g, cp = P["gamma"], P["cp"]
        lnT = P["lnT0"] + g * s / cp + (g - 1) * (lr - P["lnrho0"])
        SS = sum(S[i][j] ** 2 for i in range(3) for j in range(3))
        heat_src = (P["eta"] * P["mu0"] * (J[0] ** 2 + J[1] ** 2 + J[2] ** 2) +
                    2 * rho * P["nu"] * SS + P["zeta"] * rho * divu ** 2)
        first = g / cp * lap(hs) + (g - 1) * lap(hlr)
        second = [g / cp * gs[i] + (g - 1) * glr[i] for i in range(3)]
        third = [g * (gs[i] / cp + glr[i]) - glr[i] for i in range(3)]
        chi = P["K"] / (rho * cp)
        heat = cp * chi * (first + sum(second[i] * third[i] for i in range(3)))
        ent = -(u[0] * gs[0] + u[1] * gs[1] + u[2] * gs[2]) + heat_src / (rho * np.exp(lnT)) + heat
    else:
        ent = np.zeros_like(lr)
    return [cont] + mom + ind + [ent]
