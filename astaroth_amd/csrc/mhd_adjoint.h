// mhd_adjoint.h -- tangent-linear (JVP) and adjoint (VJP) of the per-cell right-hand side
// mhd_rhs (mhd_common.h), written by hand. There is no reference counterpart: the reference
// has no differentiation of its solver.
//
// Inputs are the derivative quantities of one cell (CellDerivs) with the upwind terms kept per
// direction (the RHS uses |u_j| up_j(f)), so that the subgradient sign(u_j) of |u_j| is exact
// (sign(0) = 0). nu and eta are differentiable parameters.
#pragma once

#include "mhd_common.h"

namespace acamd {

// Layout of the derivative adjoints: the adjoint of every derivative quantity the right-hand
// side reads, which the transposed stencils gather from the neighbours, one m^3 field each, NG
// in all. The adjoints of a cell's own values (lnrho, u, ss) need no stencil: the local stage
// adds them to lambda_f directly. The diagonal derivatives d_i A_i do not enter the right-hand
// side (B = curl A, and grad(div A) comes from the second and mixed stencils), so they have no
// field.
//   lnrho (q = 0) and u_i (q = 1..3): d_x, d_y, d_z, lap, up_x, up_y, up_z  at 7 q
//   grad(div u) x y z                                                        at G_GDU
//   A_i (q = 4..6): the two off-diagonal first derivatives (x before y before z), lap
//                                                                            at 31 + 3 (q - 4)
//   grad(div A) x y z                                                        at G_GDA
//   ss (q = 7): d_x, d_y, d_z, lap                                           at 43
enum : int { ROLE_DX = 0, ROLE_DY = 1, ROLE_DZ = 2, ROLE_LAP = 3, ROLE_UPX = 4, ROLE_UPY = 5,
             ROLE_UPZ = 6 };
constexpr int NG = 47;
constexpr int G_GDU = 28;
constexpr int G_GDA = 40;
// Field index of the adjoint of derivative `role` of field q, or -1 when there is none.
__host__ __device__ constexpr int g_index(int q, int role)
{
    if (q <= 3)
        return 7 * q + role;
    if (q <= 6) {
        const int i = q - 4;
        if (role == ROLE_LAP)
            return 31 + 3 * i + 2;
        if (role > ROLE_DZ || role == i)
            return -1;
        return 31 + 3 * i + (role < i ? role : role - 1);
    }
    return role <= ROLE_LAP ? 43 + role : -1;
}

// Upwind derivatives per direction: ul[j] = up_j(lnrho), uu[i][j] = up_j(u_i).
template <typename T>
struct UpwindDirs {
    T ul[3];
    T uu[3][3];
};

__device__ __forceinline__ float ac_sign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ double ac_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// Fills d.upl and d.upu from the per-direction upwind derivatives.
template <typename T>
__device__ __forceinline__ void combine_upwind(CellDerivs<T>& d, const UpwindDirs<T>& up)
{
    const T au[3] = {ac_abs(d.u[0]), ac_abs(d.u[1]), ac_abs(d.u[2])};
    d.upl = au[0] * up.ul[0] + au[1] * up.ul[1] + au[2] * up.ul[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        d.upu[i] = au[0] * up.uu[i][0] + au[1] * up.uu[i][1] + au[2] * up.uu[i][2];
}

// ------------------------------------------------------------------------------------------
// JVP: rhs_dot = dR/d(derivs) . dd + dR/dnu * dnu + dR/deta * deta, at the point d.
// dd holds the same derivative quantities of the tangent field; dup its upwind derivatives.
template <typename T, bool UPWD, bool ENTROPY>
__device__ __forceinline__ void mhd_rhs_jvp(const CellDerivs<T>& d, const UpwindDirs<T>& up,
                                            const CellDerivs<T>& dd, const UpwindDirs<T>& dup,
                                            const PhysParams<T>& P, T dnu, T deta, T out[NF])
{
    const T third = T(1) / T(3);
    const T divu = d.gu[0][0] + d.gu[1][1] + d.gu[2][2];
    const T ddivu = dd.gu[0][0] + dd.gu[1][1] + dd.gu[2][2];

    // upwind: upl = sum_j |u_j| ul_j
    T dupl = T(0), dupu[3] = {T(0), T(0), T(0)};
    if (UPWD) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T au = ac_abs(d.u[j]), sg = ac_sign(d.u[j]);
            dupl += au * dup.ul[j] + sg * dd.u[j] * up.ul[j];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                dupu[i] += au * dup.uu[i][j] + sg * dd.u[j] * up.uu[i][j];
        }
    }
    // continuity
    T c = -(d.u[0] * dd.glr[0] + d.u[1] * dd.glr[1] + d.u[2] * dd.glr[2]) -
          (dd.u[0] * d.glr[0] + dd.u[1] * d.glr[1] + dd.u[2] * d.glr[2]) - ddivu;
    if (UPWD)
        c += dupl;
    out[0] = c;

    T S[3][3], dS[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S[i][j] = T(0.5) * (d.gu[i][j] + d.gu[j][i]);
            dS[i][j] = T(0.5) * (dd.gu[i][j] + dd.gu[j][i]);
        }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        S[i][i] = T(2) * third * d.gu[i][i] - third * (divu - d.gu[i][i]);
        dS[i][i] = T(2) * third * dd.gu[i][i] - third * (ddivu - dd.gu[i][i]);
    }
    const T B[3] = {d.ga[2][1] - d.ga[1][2], d.ga[0][2] - d.ga[2][0], d.ga[1][0] - d.ga[0][1]};
    const T dB[3] = {dd.ga[2][1] - dd.ga[1][2], dd.ga[0][2] - dd.ga[2][0], dd.ga[1][0] - dd.ga[0][1]};
    T j[3], dj[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        j[i] = P.inv_mu0 * (d.gda[i] - d.lapa[i]);
        dj[i] = P.inv_mu0 * (dd.gda[i] - dd.lapa[i]);
    }
    const T rho = ac_exp(d.lr);
    const T inv_rho = T(1) / rho;
    const T drho = rho * dd.lr;
    const T dinv_rho = -inv_rho * dd.lr;
    T cs2 = P.cs2_sound, dcs2 = T(0), lnT_rel = T(0), dlnT_rel = T(0);
    if (ENTROPY) {
        lnT_rel = P.gamma_inv_cp * d.s + P.gm1 * (d.lr - P.lnrho0);
        dlnT_rel = P.gamma_inv_cp * dd.s + P.gm1 * dd.lr;
        cs2 = P.cs2_sound * ac_exp(lnT_rel);
        dcs2 = cs2 * dlnT_rel;
    }
    const T jxB[3] = {j[1] * B[2] - j[2] * B[1], j[2] * B[0] - j[0] * B[2], j[0] * B[1] - j[1] * B[0]};
    const T djxB[3] = {dj[1] * B[2] + j[1] * dB[2] - dj[2] * B[1] - j[2] * dB[1],
                       dj[2] * B[0] + j[2] * dB[0] - dj[0] * B[2] - j[0] * dB[2],
                       dj[0] * B[1] + j[0] * dB[1] - dj[1] * B[0] - j[1] * dB[0]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T m = -(dd.gu[i][0] * d.u[0] + dd.gu[i][1] * d.u[1] + dd.gu[i][2] * d.u[2]) -
              (d.gu[i][0] * dd.u[0] + d.gu[i][1] * dd.u[1] + d.gu[i][2] * dd.u[2]);
        if (UPWD)
            m += dupu[i];
        const T press = ENTROPY ? (P.inv_cp * d.gs[i] + d.glr[i]) : d.glr[i];
        const T dpress = ENTROPY ? (P.inv_cp * dd.gs[i] + dd.glr[i]) : dd.glr[i];
        m -= dcs2 * press + cs2 * dpress;
        m += dinv_rho * jxB[i] + inv_rho * djxB[i];
        const T Sg = S[i][0] * d.glr[0] + S[i][1] * d.glr[1] + S[i][2] * d.glr[2];
        const T dSg = dS[i][0] * d.glr[0] + dS[i][1] * d.glr[1] + dS[i][2] * d.glr[2] +
                      S[i][0] * dd.glr[0] + S[i][1] * dd.glr[1] + S[i][2] * dd.glr[2];
        const T visc = d.lapu[i] + third * d.gdu[i] + T(2) * Sg;
        const T dvisc = dd.lapu[i] + third * dd.gdu[i] + T(2) * dSg;
        m += P.nu * dvisc + dnu * visc;
        m += P.zeta * dd.gdu[i];
        out[1 + i] = m;
    }
    out[4] = (dd.u[1] * B[2] + d.u[1] * dB[2] - dd.u[2] * B[1] - d.u[2] * dB[1]) +
             P.eta * dd.lapa[0] + deta * d.lapa[0];
    out[5] = (dd.u[2] * B[0] + d.u[2] * dB[0] - dd.u[0] * B[2] - d.u[0] * dB[2]) +
             P.eta * dd.lapa[1] + deta * d.lapa[1];
    out[6] = (dd.u[0] * B[1] + d.u[0] * dB[1] - dd.u[1] * B[0] - d.u[1] * dB[0]) +
             P.eta * dd.lapa[2] + deta * d.lapa[2];
    if (ENTROPY) {
        const T inv_pT = ac_exp(-(d.lr + lnT_rel + P.lnT0));
        const T dinv_pT = -inv_pT * (dd.lr + dlnT_rel);
        T cS = T(0), dcS = T(0);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                cS += S[a][b] * S[a][b];
                dcS += T(2) * S[a][b] * dS[a][b];
            }
        const T jj = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
        const T djj = T(2) * (j[0] * dj[0] + j[1] * dj[1] + j[2] * dj[2]);
        const T Q = P.eta_mu0 * jj + T(2) * rho * P.nu * cS + P.zeta * rho * divu * divu;
        const T dQ = P.eta_mu0 * djj + deta * P.mu0 * jj + T(2) * (drho * P.nu * cS + rho * dnu * cS + rho * P.nu * dcS) +
                     P.zeta * (drho * divu * divu + rho * T(2) * divu * ddivu);
        const T first = P.gamma_inv_cp * d.laps + P.gm1 * d.laplr;
        const T dfirst = P.gamma_inv_cp * dd.laps + P.gm1 * dd.laplr;
        T st = T(0), dst = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T sec = P.gamma_inv_cp * d.gs[i] + P.gm1 * d.glr[i];
            const T thi = P.gamma * (P.inv_cp * d.gs[i] + d.glr[i]) - d.glr[i];
            const T dsec = P.gamma_inv_cp * dd.gs[i] + P.gm1 * dd.glr[i];
            const T dthi = P.gamma * (P.inv_cp * dd.gs[i] + dd.glr[i]) - dd.glr[i];
            st += sec * thi;
            dst += dsec * thi + sec * dthi;
        }
        const T heat = P.K * inv_rho * (first + st);
        const T dheat = P.K * (dinv_rho * (first + st) + inv_rho * (dfirst + dst));
        out[7] = -(dd.u[0] * d.gs[0] + dd.u[1] * d.gs[1] + dd.u[2] * d.gs[2]) -
                 (d.u[0] * dd.gs[0] + d.u[1] * dd.gs[1] + d.u[2] * dd.gs[2]) + dinv_pT * Q +
                 inv_pT * dQ + dheat;
    }
    else {
        out[7] = T(0);
    }
}

// ------------------------------------------------------------------------------------------
// VJP: a = (dR/d(derivs))^T mu, plus a_nu, a_eta. a and aup are overwritten.
template <typename T, bool UPWD, bool ENTROPY>
__device__ __forceinline__ void mhd_rhs_vjp(const CellDerivs<T>& d, const UpwindDirs<T>& up,
                                            const T mu[NF], const PhysParams<T>& P,
                                            CellDerivs<T>& a, UpwindDirs<T>& aup, T& a_nu,
                                            T& a_eta)
{
    const T third = T(1) / T(3);
    // zero the adjoint
    a.lr = a.laplr = a.upl = a.s = a.laps = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        a.glr[i] = a.u[i] = a.lapu[i] = a.gdu[i] = a.upu[i] = a.lapa[i] = a.gda[i] = a.gs[i] = T(0);
        aup.ul[i] = T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.gu[i][k] = a.ga[i][k] = T(0);
            aup.uu[i][k] = T(0);
        }
    }
    a_nu = T(0);
    a_eta = T(0);

    // forward quantities
    const T divu = d.gu[0][0] + d.gu[1][1] + d.gu[2][2];
    T S[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            S[i][k] = T(0.5) * (d.gu[i][k] + d.gu[k][i]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        S[i][i] = T(2) * third * d.gu[i][i] - third * (divu - d.gu[i][i]);
    const T B[3] = {d.ga[2][1] - d.ga[1][2], d.ga[0][2] - d.ga[2][0], d.ga[1][0] - d.ga[0][1]};
    T j[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        j[i] = P.inv_mu0 * (d.gda[i] - d.lapa[i]);
    const T rho = ac_exp(d.lr);
    const T inv_rho = T(1) / rho;
    T cs2 = P.cs2_sound, lnT_rel = T(0);
    if (ENTROPY) {
        lnT_rel = P.gamma_inv_cp * d.s + P.gm1 * (d.lr - P.lnrho0);
        cs2 = P.cs2_sound * ac_exp(lnT_rel);
    }
    const T jxB[3] = {j[1] * B[2] - j[2] * B[1], j[2] * B[0] - j[0] * B[2], j[0] * B[1] - j[1] * B[0]};

    // adjoint intermediates
    T a_divu = T(0), a_S[3][3], a_B[3] = {T(0), T(0), T(0)}, a_j[3] = {T(0), T(0), T(0)};
    T a_rho = T(0), a_inv_rho = T(0), a_cs2 = T(0), a_lnT_rel = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            a_S[i][k] = T(0);

    // entropy
    if (ENTROPY) {
        const T m7 = mu[7];
        const T inv_pT = ac_exp(-(d.lr + lnT_rel + P.lnT0));
        T cS = T(0);
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                cS += S[p][q] * S[p][q];
        const T jj = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
        const T Q = P.eta_mu0 * jj + T(2) * rho * P.nu * cS + P.zeta * rho * divu * divu;
        const T first = P.gamma_inv_cp * d.laps + P.gm1 * d.laplr;
        T sec[3], thi[3], st = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            sec[i] = P.gamma_inv_cp * d.gs[i] + P.gm1 * d.glr[i];
            thi[i] = P.gamma * (P.inv_cp * d.gs[i] + d.glr[i]) - d.glr[i];
            st += sec[i] * thi[i];
        }
        // ent = -(u . gs) + inv_pT Q + K inv_rho (first + st)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a.u[i] -= m7 * d.gs[i];
            a.gs[i] -= m7 * d.u[i];
        }
        const T a_inv_pT = m7 * Q;
        const T a_Q = m7 * inv_pT;
        a_inv_rho += m7 * P.K * (first + st);
        const T a_fs = m7 * P.K * inv_rho;  // adjoint of (first + st)
        a.laps += a_fs * P.gamma_inv_cp;
        a.laplr += a_fs * P.gm1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T a_sec = a_fs * thi[i], a_thi = a_fs * sec[i];
            a.gs[i] += a_sec * P.gamma_inv_cp + a_thi * P.gamma * P.inv_cp;
            a.glr[i] += a_sec * P.gm1 + a_thi * (P.gamma - T(1));
        }
        // Q = eta mu0 jj + 2 rho nu cS + zeta rho divu^2
#pragma unroll
        for (int i = 0; i < 3; ++i)
            a_j[i] += a_Q * P.eta_mu0 * T(2) * j[i];
        a_rho += a_Q * (T(2) * P.nu * cS + P.zeta * divu * divu);
        const T a_cS = a_Q * T(2) * rho * P.nu;
        a_divu += a_Q * P.zeta * rho * T(2) * divu;
        a_nu += a_Q * T(2) * rho * cS;
        a_eta += a_Q * P.mu0 * jj;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                a_S[p][q] += a_cS * T(2) * S[p][q];
        // inv_pT = exp(-(lr + lnT_rel + lnT0))
        a.lr -= a_inv_pT * inv_pT;
        a_lnT_rel -= a_inv_pT * inv_pT;
    }
    // induction
    {
        const T m4 = mu[4], m5 = mu[5], m6 = mu[6];
        a.u[1] += m4 * B[2];
        a_B[2] += m4 * d.u[1];
        a.u[2] -= m4 * B[1];
        a_B[1] -= m4 * d.u[2];
        a.lapa[0] += m4 * P.eta;
        a_eta += m4 * d.lapa[0];
        a.u[2] += m5 * B[0];
        a_B[0] += m5 * d.u[2];
        a.u[0] -= m5 * B[2];
        a_B[2] -= m5 * d.u[0];
        a.lapa[1] += m5 * P.eta;
        a_eta += m5 * d.lapa[1];
        a.u[0] += m6 * B[1];
        a_B[1] += m6 * d.u[0];
        a.u[1] -= m6 * B[0];
        a_B[0] -= m6 * d.u[1];
        a.lapa[2] += m6 * P.eta;
        a_eta += m6 * d.lapa[2];
    }
    // momentum
    T a_jxB[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const T m = mu[1 + i];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.gu[i][k] -= m * d.u[k];
            a.u[k] -= m * d.gu[i][k];
        }
        if (UPWD)
            a.upu[i] += m;
        const T press = ENTROPY ? (P.inv_cp * d.gs[i] + d.glr[i]) : d.glr[i];
        a_cs2 -= m * press;
        const T a_press = -m * cs2;
        if (ENTROPY)
            a.gs[i] += a_press * P.inv_cp;
        a.glr[i] += a_press;
        a_inv_rho += m * jxB[i];
        a_jxB[i] = m * inv_rho;
        const T Sg = S[i][0] * d.glr[0] + S[i][1] * d.glr[1] + S[i][2] * d.glr[2];
        a.lapu[i] += m * P.nu;
        a.gdu[i] += m * (P.nu * third + P.zeta);
        const T a_Sg = m * P.nu * T(2);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a_S[i][k] += a_Sg * d.glr[k];
            a.glr[k] += a_Sg * S[i][k];
        }
        a_nu += m * (d.lapu[i] + third * d.gdu[i] + T(2) * Sg);
    }
    // jxB
    a_j[1] += a_jxB[0] * B[2];
    a_B[2] += a_jxB[0] * j[1];
    a_j[2] -= a_jxB[0] * B[1];
    a_B[1] -= a_jxB[0] * j[2];
    a_j[2] += a_jxB[1] * B[0];
    a_B[0] += a_jxB[1] * j[2];
    a_j[0] -= a_jxB[1] * B[2];
    a_B[2] -= a_jxB[1] * j[0];
    a_j[0] += a_jxB[2] * B[1];
    a_B[1] += a_jxB[2] * j[0];
    a_j[1] -= a_jxB[2] * B[0];
    a_B[0] -= a_jxB[2] * j[1];
    // cs2 = cs2_sound exp(lnT_rel)
    if (ENTROPY) {
        a_lnT_rel += a_cs2 * cs2;
        a.s += a_lnT_rel * P.gamma_inv_cp;
        a.lr += a_lnT_rel * P.gm1;
    }
    // rho = exp(lr), inv_rho = 1 / rho
    a.lr += a_rho * rho - a_inv_rho * inv_rho;
    // j = (gda - lapa) / mu0
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        a.gda[i] += a_j[i] * P.inv_mu0;
        a.lapa[i] -= a_j[i] * P.inv_mu0;
    }
    // B = curl A
    a.ga[2][1] += a_B[0];
    a.ga[1][2] -= a_B[0];
    a.ga[0][2] += a_B[1];
    a.ga[2][0] -= a_B[1];
    a.ga[1][0] += a_B[2];
    a.ga[0][1] -= a_B[2];
    // continuity: -(u . glr) - divu + upl
    {
        const T m0 = mu[0];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a.u[i] -= m0 * d.glr[i];
            a.glr[i] -= m0 * d.u[i];
        }
        a_divu -= m0;
        if (UPWD)
            a.upl += m0;
    }
    // S (symmetric in use: both a_S[i][k] and a_S[k][i] feed S_ik)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (i == k) {
                // S_ii = gu_ii - divu / 3
                a.gu[i][i] += a_S[i][i];
                a_divu -= a_S[i][i] * third;
            }
            else {
                a.gu[i][k] += T(0.5) * a_S[i][k];
                a.gu[k][i] += T(0.5) * a_S[i][k];
            }
        }
    // divu
#pragma unroll
    for (int i = 0; i < 3; ++i)
        a.gu[i][i] += a_divu;
    // upwind: upl = sum_j |u_j| ul_j, upu_i = sum_j |u_j| uu_ij
    if (UPWD) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const T au = ac_abs(d.u[k]), sg = ac_sign(d.u[k]);
            aup.ul[k] = a.upl * au;
            T acc = a.upl * up.ul[k];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                aup.uu[i][k] = a.upu[i] * au;
                acc += a.upu[i] * up.uu[i][k];
            }
            a.u[k] += sg * acc;
        }
    }
}

// Hands the derivative adjoints of one cell to put(field index, value) in the layout above
// (the upwind ones only with upwinding: without it no transposed stencil reads them).
template <bool UPWD, typename T, typename Put>
__device__ __forceinline__ void put_derivative_adjoints(const CellDerivs<T>& ad,
                                                        const UpwindDirs<T>& aup, Put&& put)
{
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        put(g_index(0, ROLE_DX + ax), ad.glr[ax]);
        if (UPWD)
            put(g_index(0, ROLE_UPX + ax), aup.ul[ax]);
        put(g_index(7, ROLE_DX + ax), ad.gs[ax]);
    }
    put(g_index(0, ROLE_LAP), ad.laplr);
    put(g_index(7, ROLE_LAP), ad.laps);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            put(g_index(1 + i, ROLE_DX + ax), ad.gu[i][ax]);
            if (UPWD)
                put(g_index(1 + i, ROLE_UPX + ax), aup.uu[i][ax]);
            if (ax != i)
                put(g_index(4 + i, ROLE_DX + ax), ad.ga[i][ax]);
        }
        put(g_index(1 + i, ROLE_LAP), ad.lapu[i]);
        put(g_index(4 + i, ROLE_LAP), ad.lapa[i]);
        put(G_GDU + i, ad.gdu[i]);
        put(G_GDA + i, ad.gda[i]);
    }
}

// The adjoints of the cell's own values, per field (zero for A, whose values the right-hand
// side does not read).
template <typename T>
__device__ __forceinline__ T value_adjoint(const CellDerivs<T>& ad, int q)
{
    return q == 0 ? ad.lr : (q <= 3 ? ad.u[q - 1] : (q == 7 ? ad.s : T(0)));
}

}  // namespace acamd
