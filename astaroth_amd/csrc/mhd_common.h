// mhd_common.h -- device-side pieces shared by every solve kernel of the MHD hot path:
// launch arguments, the per-cell right-hand side, and the fused 2N-RK3 state update.
//
// Physics follows ref acc-runtime/samples/mhd_modular/mhdsolver.ac:279-426 (continuity,
// momentum, induction, entropy with heat conduction) and ref src/utils/modelsolver.cc:772-964.
// Stencils are the 6th-order ones of ref acc-runtime/stdlib/derivs.h:374-427,591-752.
// The RK3 coefficients are Williamson's 2N set, ref acc-runtime/stdlib/integrators.h:11-12.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace acamd {

constexpr int NF = 8;  // lnrho, ux, uy, uz, ax, ay, az, ss
constexpr int R = 3;   // stencil radius

// Scalar parameters of the right-hand side, in the compute type T.
template <typename T>
struct PhysParams {
    T cs2_sound, nu, zeta, eta, mu0, inv_mu0, cp, inv_cp, gamma, gm1, lnT0, lnrho0, K;
    T gamma_inv_cp;  // gamma / cp
    T eta_mu0;       // eta * mu0
    T nu_third;      // nu / 3
    T pT_scale;      // cs2_sound * exp(-lnT0): 1 / (rho T) = pT_scale / (rho cs2)
};

// Stencil coefficients already multiplied by the grid spacing factors, per axis (x, y, z).
// d1[a][o]: first derivative, offset o = 1..3 (antisymmetric), d1[a][0] unused.
// d2[a][o]: second derivative, o = 0 is the centre weight.
// up[a][o]: upwind (hyperdiffusion) operator, o = 0 is the centre weight.
// mx[p][o]: mixed derivative along the diagonals of plane p (0 = xy, 1 = xz, 2 = yz).
template <typename T>
struct Coeffs {
    T d1[3][4];
    T d2[3][4];
    T up[3][4];
    T mx[3][4];
};

// Everything one fused substep launch needs. Passed by value (well under the 4 KiB limit).
template <typename T>
struct SolveArgs {
    const T* in[NF];   // state f_s, halos filled
    T* out[NF];        // f_{s+1}, interior written
    T* w[NF];          // 2N register, updated in place
    void* comp[NF];    // Kahan compensation (FP32 build, optional), updated in place
    int nx, ny, nz;    // interior extent
    int mx, my, mz;    // extent with halos
    long long sz;      // mx * my
    // Elements between consecutive fields of in, out and w (each role is one slab; the tiled
    // kernels address field f as slab + f * fstride and require that layout)
    long long fstride;
    Coeffs<T> C;
    PhysParams<T> P;
    T dt;              // used when dt_dev == nullptr
    T alpha, beta;     // alpha_s, beta_s
    int step;          // substep 0, 1, 2
    // Device-resident timestep (umax2_cur == nullptr: use dt). umax2_cur / umax2_next hold a
    // pair (x, y) of reduced maxima (double bits); what they hold and the dt formula depend on
    // cfl_mode (see cfl_cell / solve_dt).
    const double* umax2_cur;
    unsigned long long* umax2_next;  // step 2: atomic max of the pair of the state reduced
    double* time_acc;                // step 2: accumulated simulated time
    double cfl_cdt, cfl_dsmin, cfl_cs;
    int cfl_mode;                    // 0 courant, 1 reference, 2 reference additive
    double cfl_cs2, cfl_cdtv, cfl_visc;  // cs^2, cdtv, max(nu, eta[, gamma chi])
    // 1: the tiled (delay-line) kernels read the x / y / z halo of `in` from the periodic images
    // inside the interior, so the halos need not be filled; 0: they read the halos as stored
    int wrap;
};

__device__ __forceinline__ float ac_exp(float x) { return expf(x); }
__device__ __forceinline__ double ac_exp(double x) { return exp(x); }
__device__ __forceinline__ float ac_abs(float x) { return fabsf(x); }
__device__ __forceinline__ double ac_abs(double x) { return fabs(x); }
// 1 / x. FP32: the hardware reciprocal estimate refined by one Newton step
// r' = r + r (1 - x r), within an ulp of the quotient for the positive, normal arguments it gets
// here (exponentials of the state), in 4 VALU instructions instead of the IEEE division's 10.
// FP64: the division.
__device__ __forceinline__ float ac_rcp(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double ac_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float ac_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double ac_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Derivative quantities of one cell that the right-hand side depends on.
// gu[i][j] = d_j u_i, ga[i][j] = d_j A_i. upl = |u| . upwind-grad(lnrho),
// upu[i] = sum_j upwind_j(u_i) |u_j| (ref modelsolver.cc:776-779,843-846).
template <typename T>
struct CellDerivs {
    T lr, glr[3], laplr, upl;
    T u[3], gu[3][3], lapu[3], gdu[3], upu[3];
    T ga[3][3], lapa[3], gda[3];
    T s, gs[3], laps;
};

// Timestep modes. 0: dt = cdt dsmin / (cs + max|u|), with max|u| of the new state, reduced by
// the third substep's solve kernel (x = |u|^2).
// 1: the reference's default CFL (ref samples/standalone_mpi/main.cc:395-440):
//    dt = min(cdt dsmin / (max|u| + sqrt(cs^2 + max vA^2)), cdtv dsmin^2 / max(nu, eta)),
//    x = |u|^2, y = vA^2 = |curl A|^2 / rho of the state entering the third substep, where the
//    reference reduces them (ref mhdsolver.ac:551-583,611).
// 2: the reference's additive CFL (AC_additive_timestep): dt = min over cells of
//    cdt dsmin / (|u| + sqrt(cs^2 + vA^2)) and cdtv dsmin^2 / max(nu, eta, gamma chi),
//    x = (|u| + sqrt(cs^2 + vA^2))^2.
// cfl_cell gives (x, y) of one cell for modes 1 and 2 from its derivative quantities.
template <typename T>
__device__ __forceinline__ void cfl_cell(int mode, double cs2, const CellDerivs<T>& d, double& x,
                                         double& y)
{
    const double u2 = (double)d.u[0] * d.u[0] + (double)d.u[1] * d.u[1] + (double)d.u[2] * d.u[2];
    const double b0 = (double)d.ga[2][1] - d.ga[1][2];
    const double b1 = (double)d.ga[0][2] - d.ga[2][0];
    const double b2 = (double)d.ga[1][0] - d.ga[0][1];
    const double va2 = (b0 * b0 + b1 * b1 + b2 * b2) / exp((double)d.lr);
    if (mode == 1) {
        x = u2;
        y = va2;
    }
    else {
        const double sm = sqrt(u2) + sqrt(cs2 + va2);
        x = sm * sm;
        y = 0.0;
    }
}

__host__ __device__ __forceinline__ double cfl_dt(int mode, double x, double y, double cdt,
                                                  double dsmin, double cs, double cs2, double cdtv,
                                                  double visc)
{
    if (mode == 0)
        return cdt * dsmin / (cs + sqrt(x));
    const double adv = mode == 1 ? cdt * dsmin / (sqrt(x) + sqrt(cs2 + y)) : cdt * dsmin / sqrt(x);
    const double dif = visc > 0.0 ? cdtv * dsmin * dsmin / visc : adv;
    return adv < dif ? adv : dif;
}

// Right-hand side of the eight evolved fields at one cell. EXACT (the precise FP32 mode): 1/rho
// by the correctly rounded division and 1/(rho T) by its own exponential, as in FP64, instead of
// the reciprocal with one Newton step and 1/(rho T) from the exponentials already taken.
template <typename T, bool UPWD, bool ENTROPY, bool EXACT = false>
__device__ __forceinline__ void mhd_rhs(const CellDerivs<T>& d, const PhysParams<T>& P, T rhs[NF])
{
    const T divu = d.gu[0][0] + d.gu[1][1] + d.gu[2][2];
    // continuity (ref mhdsolver.ac:334-340)
    T cont = -(d.u[0] * d.glr[0] + d.u[1] * d.glr[1] + d.u[2] * d.glr[2]) - divu;
    if (UPWD)
        cont += d.upl;
    rhs[0] = cont;

    // traceless rate of strain (ref operators.h:145-161)
    const T third = T(1) / T(3);
    const T S00 = T(2) * third * d.gu[0][0] - third * (d.gu[1][1] + d.gu[2][2]);
    const T S11 = T(2) * third * d.gu[1][1] - third * (d.gu[0][0] + d.gu[2][2]);
    const T S22 = T(2) * third * d.gu[2][2] - third * (d.gu[0][0] + d.gu[1][1]);
    const T S01 = T(0.5) * (d.gu[0][1] + d.gu[1][0]);
    const T S02 = T(0.5) * (d.gu[0][2] + d.gu[2][0]);
    const T S12 = T(0.5) * (d.gu[1][2] + d.gu[2][1]);

    // magnetic field and current (ref mhdsolver.ac:354-358)
    const T B0 = d.ga[2][1] - d.ga[1][2];
    const T B1 = d.ga[0][2] - d.ga[2][0];
    const T B2 = d.ga[1][0] - d.ga[0][1];
    const T j0 = P.inv_mu0 * (d.gda[0] - d.lapa[0]);
    const T j1 = P.inv_mu0 * (d.gda[1] - d.lapa[1]);
    const T j2 = P.inv_mu0 * (d.gda[2] - d.lapa[2]);

    const T rho = ac_exp(d.lr);
    const T inv_rho = EXACT ? T(1) / rho : ac_rcp(rho);

    T cs2;
    T lnT_rel = T(0);  // lnT - lnT0
    if (ENTROPY) {
        lnT_rel = P.gamma_inv_cp * d.s + P.gm1 * (d.lr - P.lnrho0);
        cs2 = P.cs2_sound * ac_exp(lnT_rel);
    }
    else {
        cs2 = P.cs2_sound;
    }

    const T jxB0 = j1 * B2 - j2 * B1;
    const T jxB1 = j2 * B0 - j0 * B2;
    const T jxB2 = j0 * B1 - j1 * B0;
    const T jxB[3] = {jxB0, jxB1, jxB2};
    const T Srow[3][3] = {{S00, S01, S02}, {S01, S11, S12}, {S02, S12, S22}};

    // momentum (ref mhdsolver.ac:342-384)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T m = -(d.gu[i][0] * d.u[0] + d.gu[i][1] * d.u[1] + d.gu[i][2] * d.u[2]);
        if (UPWD)
            m += d.upu[i];
        const T press = ENTROPY ? (P.inv_cp * d.gs[i] + d.glr[i]) : d.glr[i];
        m -= cs2 * press;
        m += inv_rho * jxB[i];
        const T Sg = Srow[i][0] * d.glr[0] + Srow[i][1] * d.glr[1] + Srow[i][2] * d.glr[2];
        m += P.nu * (d.lapu[i] + third * d.gdu[i] + T(2) * Sg);
        m += P.zeta * d.gdu[i];
        rhs[1 + i] = m;
    }

    // induction (ref mhdsolver.ac:279-281): u x B + eta lap A
    rhs[4] = (d.u[1] * B2 - d.u[2] * B1) + P.eta * d.lapa[0];
    rhs[5] = (d.u[2] * B0 - d.u[0] * B2) + P.eta * d.lapa[1];
    rhs[6] = (d.u[0] * B1 - d.u[1] * B0) + P.eta * d.lapa[2];

    // entropy (ref mhdsolver.ac:387-426)
    if (ENTROPY) {
        // 1 / (rho T) = exp(-(lnrho + lnT)); in FP32 from the exponentials already taken:
        // T = T0 exp(lnT_rel) and cs2 = cs2_sound exp(lnT_rel), so 1 / (rho T) =
        // pT_scale / (rho cs2) with pT_scale = cs2_sound / T0
        T inv_pT;
        if constexpr (sizeof(T) == 4 && !EXACT)
            inv_pT = inv_rho * P.pT_scale * ac_rcp(cs2);
        else
            inv_pT = ac_exp(-(d.lr + lnT_rel + P.lnT0));
        const T contractS = S00 * S00 + S11 * S11 + S22 * S22 +
                            T(2) * (S01 * S01 + S02 * S02 + S12 * S12);
        const T RHS = P.eta_mu0 * (j0 * j0 + j1 * j1 + j2 * j2) + T(2) * rho * P.nu * contractS +
                      P.zeta * rho * divu * divu;
        const T first = P.gamma_inv_cp * d.laps + P.gm1 * d.laplr;
        T st = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T second = P.gamma_inv_cp * d.gs[i] + P.gm1 * d.glr[i];
            const T thirdv = P.gamma * (P.inv_cp * d.gs[i] + d.glr[i]) - d.glr[i];
            st += second * thirdv;
        }
        // cp * chi with chi = K / (rho cp)
        const T heat = P.K * inv_rho * (first + st);
        rhs[7] = -(d.u[0] * d.gs[0] + d.u[1] * d.gs[1] + d.u[2] * d.gs[2]) + inv_pT * RHS + heat;
    }
    else {
        rhs[7] = T(0);
    }
}

// ---- Kahan compensation storage (FP32 build) --------------------------------------------
// The compensation c of a value f obeys |c| <= ulp(f)/2. It is stored either as a float
// (KB = 32) or as a signed integer in units of ulp(f) * 2^-KB (KB = 8 or 16), so one extra
// byte per value keeps eight more bits of the state.
__device__ __forceinline__ int ulp_exp_f32(float f)
{
    const uint32_t e = (__float_as_uint(f) >> 23) & 0xffu;
    return (int)(e > 0u ? e : 1u) - 150;  // ulp(f) = 2^ulp_exp
}

// The compensation array holds one record per cell: the NF fields' compensations of the cell
// side by side (NF * KB / 8 bytes), so the tiled kernels move a cell's record in one or two
// vector accesses. Field q's value of cell c is element c * NF of the array that starts at
// byte q * KB / 8 (`base` below).
template <int KB>
__device__ __forceinline__ float comp_load(const void* base, long long c, float f)
{
    if (KB == 32) {
        return ((const float*)base)[c * NF];
    }
    else if (KB == 16) {
        const int q = ((const int16_t*)base)[c * NF];
        return ldexpf((float)q, ulp_exp_f32(f) - 16);
    }
    else {
        const int q = ((const int8_t*)base)[c * NF];
        return ldexpf((float)q, ulp_exp_f32(f) - 8);
    }
}

template <int KB>
__device__ __forceinline__ void comp_store(void* base, long long c, float comp, float f)
{
    if (KB == 32) {
        ((float*)base)[c * NF] = comp;
    }
    else if (KB == 16) {
        float q = rintf(ldexpf(comp, 16 - ulp_exp_f32(f)));
        q = fminf(fmaxf(q, -32767.f), 32767.f);
        ((int16_t*)base)[c * NF] = (int16_t)q;
    }
    else {
        float q = rintf(ldexpf(comp, 8 - ulp_exp_f32(f)));
        q = fminf(fmaxf(q, -127.f), 127.f);
        ((int8_t*)base)[c * NF] = (int8_t)q;
    }
}

__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }

// Fused 2N-RK3 update of one value (ref integrators.h:87-101, modelsolver.cc:1126-1159):
//   w' = alpha_s w + dt rhs        (w is not read at s = 0: alpha_0 = 0)
//   f' = f + beta_s w'             (plain: rounded product, then rounded sum, like the
//                                   reference's final pass; KB > 0: Kahan-compensated)
// w' is stored for s < 2 only (it is dead after the last substep).
template <typename T, int KB>
__device__ __forceinline__ void rk_update(const SolveArgs<T>& a, int q, long long c, T f, T rhs,
                                          T dt, T* fout)
{
    T wn = dt * rhs;
    if (a.step > 0)
        wn = a.alpha * a.w[q][c] + wn;
    if (a.step < 2)
        a.w[q][c] = wn;
    if (KB == 0) {
        *fout = add_rn(f, mul_rn(a.beta, wn));
    }
    else {
        const float cc = comp_load<KB>(a.comp[q], c, f);
        const float y = a.beta * wn - cc;
        const float t = f + y;
        const float cn = (t - f) - y;
        comp_store<KB>(a.comp[q], c, cn, t);
        *fout = t;
    }
}

// Access at (uniform base, 32-bit byte offset): the form global loads and stores take as one
// scalar base register pair plus one vector offset, with no per-lane 64-bit address math.
template <typename T>
__device__ __forceinline__ T ld_b(const T* base, unsigned boff)
{
    return *(const T*)((const char*)base + boff);
}
template <typename T>
__device__ __forceinline__ void st_b(T* base, unsigned boff, T v)
{
    *(T*)((char*)base + boff) = v;
}
// The same through a base pointer known to be in global memory (see sbase).
typedef __attribute__((address_space(1))) char gchar;
template <typename T>
__device__ __forceinline__ T ld_g(const gchar* base, unsigned boff)
{
    return *(const __attribute__((address_space(1))) T*)(base + boff);
}
template <typename T>
__device__ __forceinline__ void st_b(gchar* base, unsigned boff, T v)
{
    *(__attribute__((address_space(1))) T*)(base + boff) = v;
}

// A cell's compensation record in registers: NF * KB / 32 dwords (2, 4 or 8).
template <int KB>
struct CompRec {
    uint32_t w[KB > 0 ? NF * KB / 32 : 1];
};
template <int KB>
__device__ __forceinline__ CompRec<KB> comp_rec_load(const char* base, unsigned boff)
{
    CompRec<KB> r;
    const char* p = base + boff;
    if constexpr (KB == 8) {
        const uint2 v = *(const uint2*)p;
        r.w[0] = v.x;
        r.w[1] = v.y;
    }
    else if constexpr (KB >= 16) {
#pragma unroll
        for (int h = 0; h < KB / 16; ++h) {
            const uint4 v = *(const uint4*)(p + 16 * h);
            r.w[4 * h + 0] = v.x;
            r.w[4 * h + 1] = v.y;
            r.w[4 * h + 2] = v.z;
            r.w[4 * h + 3] = v.w;
        }
    }
    return r;
}
template <int KB>
__device__ __forceinline__ void comp_rec_store(char* base, unsigned boff, const CompRec<KB>& r)
{
    char* p = base + boff;
    if constexpr (KB == 8) {
        *(uint2*)p = make_uint2(r.w[0], r.w[1]);
    }
    else if constexpr (KB >= 16) {
#pragma unroll
        for (int h = 0; h < KB / 16; ++h)
            *(uint4*)(p + 16 * h) =
                make_uint4(r.w[4 * h + 0], r.w[4 * h + 1], r.w[4 * h + 2], r.w[4 * h + 3]);
    }
}
// Field f's raw compensation (float bits, or the sign-extended int8 / int16).
template <int KB>
__device__ __forceinline__ int comp_rec_get(const CompRec<KB>& r, int f)
{
    if constexpr (KB == 32)
        return (int)r.w[f];
    else if constexpr (KB == 16)
        return (int)(r.w[f >> 1] << (16 - 16 * (f & 1))) >> 16;
    else if constexpr (KB == 8)
        return (int)(r.w[f >> 2] << (24 - 8 * (f & 3))) >> 24;
    else
        return 0;
}
template <int KB>
__device__ __forceinline__ void comp_rec_set(CompRec<KB>& r, int f, int raw)
{
    if constexpr (KB == 32) {
        r.w[f] = (uint32_t)raw;
    }
    else if constexpr (KB == 16) {
        const int s = 16 * (f & 1);
        r.w[f >> 1] = (r.w[f >> 1] & ~(0xffffu << s)) | (((uint32_t)raw & 0xffffu) << s);
    }
    else if constexpr (KB == 8) {
        const int s = 8 * (f & 3);
        r.w[f >> 2] = (r.w[f >> 2] & ~(0xffu << s)) | (((uint32_t)raw & 0xffu) << s);
    }
}
// Raw encoding of a compensation `comp` of the value f (the inverse of comp_decode).
template <int KB>
__device__ __forceinline__ int comp_encode(float comp, float f)
{
    if constexpr (KB == 32) {
        return __float_as_int(comp);
    }
    else {
        constexpr float lim = KB == 16 ? 32767.f : 127.f;
        float q = rintf(ldexpf(comp, KB - ulp_exp_f32(f)));
        q = fminf(fmaxf(q, -lim), lim);
        return (int)q;
    }
}
template <int KB>
__device__ __forceinline__ float comp_decode(int raw, float f)
{
    if (KB == 32)
        return __int_as_float(raw);
    return ldexpf((float)raw, ulp_exp_f32(f) - KB);
}

// The fused update addressed as (uniform plane base, 32-bit in-plane byte offset boff of the
// cell in the plane of w and f), with w of the cell (wold; ignored at step 0) and, for KB > 0,
// the raw compensation (craw) already loaded, so their latency hides behind the stencil work;
// the new compensation's raw encoding is returned in *cenc (the caller stores the cell's
// record).
template <typename T, int KB, typename P>
__device__ __forceinline__ T rk_update_pre(P __restrict__ wp, P __restrict__ op, unsigned boff,
                                           int step, T alpha, T beta, T f, T rhs, T dt, T wold,
                                           int craw, int* cenc)
{
    T wn = dt * rhs;
    if (step > 0)
        wn = alpha * wold + wn;
    if (step < 2)
        st_b(wp, boff, wn);
    T fn;
    if (KB == 0) {
        fn = add_rn(f, mul_rn(beta, wn));
    }
    else {
        const float cc = comp_decode<KB>(craw, f);
        const float y = beta * wn - cc;
        const float t = f + y;
        const float cn = (t - f) - y;
        *cenc = comp_encode<KB>(cn, t);
        fn = t;
    }
    st_b(op, boff, fn);
    return fn;
}

template <typename T>
__device__ __forceinline__ T solve_dt(const SolveArgs<T>& a)
{
    if (a.umax2_cur == nullptr)
        return a.dt;
    return (T)cfl_dt(a.cfl_mode, a.umax2_cur[0], a.umax2_cur[1], a.cfl_cdt, a.cfl_dsmin, a.cfl_cs,
                     a.cfl_cs2, a.cfl_cdtv, a.cfl_visc);
}

// Block-wide max of |u|^2 into umax2_next (step 2 of the device-dt path), plus time.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* addr, double v)
{
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

}  // namespace acamd
