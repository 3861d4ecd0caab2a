/*
 * oracle.cc -- TEST INFRASTRUCTURE ONLY. CPU restatement of the reference's host model solver
 * for the MHD hot path, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker. Nothing in astaroth_amd/ links or calls this file.
 *
 * What it restates (reference tree, file:line):
 *   long double arithmetic ............ src/utils/modelsolver.cc:68-83
 *   6th-order pencils ................. src/utils/modelsolver.cc:190-299 (first, second, cross)
 *   der* gatherers, upwind ............ src/utils/modelsolver.cc:301-487
 *   ScalarData (value/grad/hess/upw) .. src/utils/modelsolver.cc:521-552
 *   operators ......................... src/utils/modelsolver.cc:628-742
 *   continuity/momentum/induction/
 *     lnT/heat_conduction/entropy ..... src/utils/modelsolver.cc:772-964
 *   alpha / beta steps ................ src/utils/modelsolver.cc:1078-1175
 *   driver (BC before each substep) ... src/utils/modelsolver.cc:1214-1272
 *   host periodic fill ................ src/utils/memory.cc:75-149
 *   host reductions (long double) ..... src/utils/modelreduce.cc:108-340
 *   "ref-cpu" two-pass FP64 loop ...... acc-runtime/samples/mhd_modular/mhdsolver.ac:586-636 with
 *                                       stdlib/integrators.h:87-101, stdlib/derivs.h:591-752,
 *                                       executed as host loops like the CPU build
 *                                       (acc-runtime/acc/stencilgen.c:441-470), OpenMP over z.
 *
 * Storage precision follows the reference's AcReal (float or double): intermediate w and the
 * state f are rounded to storage after every alpha and beta step, and the beta step adds in
 * storage precision, exactly as solve_beta_step does.
 *
 * Parity pins (tests/test_oracle.py): polynomial exactness and Fourier symbols of every stencil,
 * RK3 amplification 1+z+z^2/2+z^3/6, Beltrami A-field decay, viscous shear-wave decay, uniform
 * state invariance, brute-force periodic fill, reductions on closed forms, and agreement of
 * ref-cpu (FP64) with the long-double path.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
typedef struct {
    int nx, ny, nz;
    double ds[3];
    double cs_sound, nu_visc, zeta, eta, mu0, cp_sound, gamma, lnT0, lnrho0, K_heatcond;
    int lupwd, lentropy;
} OracleConfig;
}

namespace {

typedef long double Scalar;
struct V3 {
    Scalar x, y, z;
};
struct M3 {
    V3 row[3];
};

static inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(const V3& a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(const Scalar s, const V3& b) { return {s * b.x, s * b.y, s * b.z}; }
static inline Scalar dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 mul(const M3& m, const V3& v)
{
    return {dot(m.row[0], v), dot(m.row[1], v), dot(m.row[2], v)};
}
static inline V3 cross(const V3& a, const V3& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Grid {
    int nx, ny, nz, mx, my, mz;
    long sy, sz;
    Scalar inv_ds[3];
    long idx(int i, int j, int k) const { return i + sy * j + sz * k; }
};

static Grid make_grid(const OracleConfig* c)
{
    Grid g;
    g.nx = c->nx;
    g.ny = c->ny;
    g.nz = c->nz;
    g.mx = c->nx + 6;
    g.my = c->ny + 6;
    g.mz = c->nz + 6;
    g.sy = g.mx;
    g.sz = (long)g.mx * g.my;
    for (int d = 0; d < 3; ++d)
        g.inv_ds[d] = (Scalar)1 / (Scalar)c->ds[d];
    return g;
}

/* 6th-order pencil operators on a 7-point pencil p[0..6], centre p[3]
 * (ref modelsolver.cc:190-299). */
static inline Scalar pencil_first(const Scalar* p, Scalar inv)
{
    const Scalar c[4] = {0, (Scalar)3 / (Scalar)4, (Scalar)-3 / (Scalar)20, (Scalar)1 / (Scalar)60};
    Scalar r = 0;
    for (int i = 1; i <= 3; ++i)
        r += c[i] * (p[3 + i] - p[3 - i]);
    return r * inv;
}
static inline Scalar pencil_second(const Scalar* p, Scalar inv)
{
    const Scalar c[4] = {(Scalar)-49 / (Scalar)18, (Scalar)3 / (Scalar)2, (Scalar)-3 / (Scalar)20,
                         (Scalar)1 / (Scalar)90};
    Scalar r = c[0] * p[3];
    for (int i = 1; i <= 3; ++i)
        r += c[i] * (p[3 + i] + p[3 - i]);
    return r * inv * inv;
}
/* Mixed derivative on the two diagonals: a = f(+i,+i) pencil, b = f(+i,-i) pencil. */
static inline Scalar pencil_cross(const Scalar* a, const Scalar* b, Scalar inva, Scalar invb)
{
    const Scalar fac = (Scalar)1 / (Scalar)720;
    const Scalar c[4] = {0, (Scalar)270 * fac, (Scalar)-27 * fac, (Scalar)2 * fac};
    Scalar r = 0;
    for (int i = 1; i <= 3; ++i)
        r += c[i] * (a[3 + i] + a[3 - i] - b[3 + i] - b[3 - i]);
    return r * inva * invb;
}
/* 6th-order upwind (hyperdiffusive) operator (ref modelsolver.cc:419-487). */
static inline Scalar pencil_upwind(const Scalar* p, Scalar inv)
{
    return (Scalar)(1.0 / 60.0) * inv *
           (-(Scalar)20 * p[3] + (Scalar)15 * (p[4] + p[2]) - (Scalar)6 * (p[5] + p[1]) + p[6] +
            p[0]);
}

struct ScalarData {
    Scalar value;
    V3 grad;
    M3 hess;
    V3 upw;
};
struct VectorData {
    ScalarData x, y, z;
};

template <typename T>
static ScalarData read_scal(const T* a, const Grid& g, int i, int j, int k)
{
    const long c = g.idx(i, j, k);
    const long s[3] = {1, g.sy, g.sz};
    ScalarData d;
    d.value = (Scalar)a[c];
    Scalar grad[3], sec[3], upw[3];
    for (int ax = 0; ax < 3; ++ax) {
        Scalar p[7];
        for (int o = 0; o < 7; ++o)
            p[o] = (Scalar)a[c + (o - 3) * s[ax]];
        grad[ax] = pencil_first(p, g.inv_ds[ax]);
        sec[ax] = pencil_second(p, g.inv_ds[ax]);
        upw[ax] = pencil_upwind(p, g.inv_ds[ax]);
    }
    Scalar mixed[3][3] = {{0}};
    for (int aa = 0; aa < 3; ++aa) {
        for (int bb = aa + 1; bb < 3; ++bb) {
            Scalar pa[7], pb[7];
            for (int o = 0; o < 7; ++o) {
                pa[o] = (Scalar)a[c + (o - 3) * s[aa] + (o - 3) * s[bb]];
                pb[o] = (Scalar)a[c + (o - 3) * s[aa] + (3 - o) * s[bb]];
            }
            mixed[aa][bb] = mixed[bb][aa] = pencil_cross(pa, pb, g.inv_ds[aa], g.inv_ds[bb]);
        }
    }
    d.grad = {grad[0], grad[1], grad[2]};
    d.upw = {upw[0], upw[1], upw[2]};
    d.hess.row[0] = {sec[0], mixed[0][1], mixed[0][2]};
    d.hess.row[1] = {mixed[1][0], sec[1], mixed[1][2]};
    d.hess.row[2] = {mixed[2][0], mixed[2][1], sec[2]};
    return d;
}

static inline Scalar laplace(const ScalarData& d)
{
    return d.hess.row[0].x + d.hess.row[1].y + d.hess.row[2].z;
}
static inline V3 vecvalue(const VectorData& v) { return {v.x.value, v.y.value, v.z.value}; }
static inline V3 vecabs(const VectorData& v)
{
    return {fabsl(v.x.value), fabsl(v.y.value), fabsl(v.z.value)};
}
static inline M3 gradients(const VectorData& v)
{
    M3 m;
    m.row[0] = v.x.grad;
    m.row[1] = v.y.grad;
    m.row[2] = v.z.grad;
    return m;
}
static inline M3 gradients_upwd(const VectorData& v)
{
    M3 m;
    m.row[0] = v.x.upw;
    m.row[1] = v.y.upw;
    m.row[2] = v.z.upw;
    return m;
}
static inline Scalar divergence(const VectorData& v) { return v.x.grad.x + v.y.grad.y + v.z.grad.z; }
static inline V3 laplace_vec(const VectorData& v)
{
    return {laplace(v.x), laplace(v.y), laplace(v.z)};
}
static inline V3 curl(const VectorData& v)
{
    return {v.z.grad.y - v.y.grad.z, v.x.grad.z - v.z.grad.x, v.y.grad.x - v.x.grad.y};
}
static inline V3 grad_div(const VectorData& v)
{
    return {v.x.hess.row[0].x + v.y.hess.row[0].y + v.z.hess.row[0].z,
            v.x.hess.row[1].x + v.y.hess.row[1].y + v.z.hess.row[1].z,
            v.x.hess.row[2].x + v.y.hess.row[2].y + v.z.hess.row[2].z};
}
/* Traceless rate of strain (ref modelsolver.cc:701-727). */
static inline M3 strain(const VectorData& v)
{
    M3 S;
    const Scalar third = (Scalar)1 / (Scalar)3, two3 = (Scalar)2 / (Scalar)3, half = (Scalar)0.5L;
    S.row[0].x = two3 * v.x.grad.x - third * (v.y.grad.y + v.z.grad.z);
    S.row[0].y = half * (v.x.grad.y + v.y.grad.x);
    S.row[0].z = half * (v.x.grad.z + v.z.grad.x);
    S.row[1].y = two3 * v.y.grad.y - third * (v.x.grad.x + v.z.grad.z);
    S.row[1].z = half * (v.y.grad.z + v.z.grad.y);
    S.row[2].z = two3 * v.z.grad.z - third * (v.x.grad.x + v.y.grad.y);
    S.row[1].x = S.row[0].y;
    S.row[2].x = S.row[0].z;
    S.row[2].y = S.row[1].z;
    return S;
}
static inline Scalar contract(const M3& m)
{
    return dot(m.row[0], m.row[0]) + dot(m.row[1], m.row[1]) + dot(m.row[2], m.row[2]);
}

struct Phys {
    Scalar cs_sound, nu, zeta, eta, mu0, cp, gamma, lnT0, lnrho0, K;
    bool upwd, entropy;
};

static Phys make_phys(const OracleConfig* c)
{
    Phys p;
    p.cs_sound = c->cs_sound;
    p.nu = c->nu_visc;
    p.zeta = c->zeta;
    p.eta = c->eta;
    p.mu0 = c->mu0;
    p.cp = c->cp_sound;
    p.gamma = c->gamma;
    p.lnT0 = c->lnT0;
    p.lnrho0 = c->lnrho0;
    p.K = c->K_heatcond;
    p.upwd = c->lupwd != 0;
    p.entropy = c->lentropy != 0;
    return p;
}

/* ref modelsolver.cc:772-781 */
static Scalar continuity(const Phys& P, const VectorData& uu, const ScalarData& lnrho)
{
    Scalar r = -dot(vecvalue(uu), lnrho.grad);
    if (P.upwd)
        r += dot(vecabs(uu), lnrho.upw);
    return r - divergence(uu);
}

/* ref modelsolver.cc:805-875 (both the LENTROPY and the isothermal branch) */
static V3 momentum(const Phys& P, const VectorData& uu, const ScalarData& lnrho,
                   const ScalarData& ss, const VectorData& aa)
{
    const M3 S = strain(uu);
    const Scalar cs2_sound = P.cs_sound * P.cs_sound;
    const Scalar cs2 = P.entropy ? cs2_sound * expl(P.gamma * ss.value / P.cp +
                                                    (P.gamma - 1) * (lnrho.value - P.lnrho0))
                                 : cs2_sound;
    const V3 j = ((Scalar)1 / P.mu0) * (grad_div(aa) - laplace_vec(aa));
    const V3 B = curl(aa);
    const Scalar inv_rho = (Scalar)1 / expl(lnrho.value);
    V3 press = P.entropy ? ((Scalar)1 / P.cp) * ss.grad + lnrho.grad : lnrho.grad;
    V3 mom = -mul(gradients(uu), vecvalue(uu)) - cs2 * press;
    if (P.upwd)
        mom = mom + mul(gradients_upwd(uu), vecabs(uu));
    mom = mom + inv_rho * cross(j, B) +
          P.nu * (laplace_vec(uu) + ((Scalar)1 / (Scalar)3) * grad_div(uu) +
                  (Scalar)2 * mul(S, lnrho.grad)) +
          P.zeta * grad_div(uu);
    return mom;
}

/* ref modelsolver.cc:877-895 */
static V3 induction(const Phys& P, const VectorData& uu, const VectorData& aa)
{
    return cross(vecvalue(uu), curl(aa)) + P.eta * laplace_vec(aa);
}

/* ref modelsolver.cc:897-964 */
static Scalar entropy(const Phys& P, const ScalarData& ss, const VectorData& uu,
                      const ScalarData& lnrho, const VectorData& aa)
{
    const M3 S = strain(uu);
    const Scalar lnT = P.lnT0 + P.gamma * ss.value / P.cp + (P.gamma - 1) * (lnrho.value - P.lnrho0);
    const Scalar inv_pT = (Scalar)1 / (expl(lnrho.value) * expl(lnT));
    const V3 j = ((Scalar)1 / P.mu0) * (grad_div(aa) - laplace_vec(aa));
    const Scalar divu = divergence(uu);
    const Scalar RHS = P.eta * P.mu0 * dot(j, j) + (Scalar)2 * expl(lnrho.value) * P.nu * contract(S) +
                       P.zeta * expl(lnrho.value) * divu * divu;
    const Scalar inv_cp = (Scalar)1 / P.cp;
    const V3 grad_ln_chi = -lnrho.grad;
    const Scalar first = P.gamma * inv_cp * laplace(ss) + (P.gamma - 1) * laplace(lnrho);
    const V3 second = (P.gamma * inv_cp) * ss.grad + (P.gamma - 1) * lnrho.grad;
    const V3 third = P.gamma * (inv_cp * ss.grad + lnrho.grad) + grad_ln_chi;
    const Scalar chi = P.K / (expl(lnrho.value) * P.cp);
    const Scalar heat = P.cp * chi * (first + dot(second, third));
    return -dot(vecvalue(uu), ss.grad) + inv_pT * RHS + heat;
}

template <typename T>
static void rhs_at(T* const* f, const Grid& g, const Phys& P, int i, int j, int k, Scalar out[8])
{
    const ScalarData lnrho = read_scal(f[0], g, i, j, k);
    VectorData uu{read_scal(f[1], g, i, j, k), read_scal(f[2], g, i, j, k),
                  read_scal(f[3], g, i, j, k)};
    VectorData aa{read_scal(f[4], g, i, j, k), read_scal(f[5], g, i, j, k),
                  read_scal(f[6], g, i, j, k)};
    const ScalarData ss = read_scal(f[7], g, i, j, k);
    out[0] = continuity(P, uu, lnrho);
    const V3 m = momentum(P, uu, lnrho, ss, aa);
    out[1] = m.x;
    out[2] = m.y;
    out[3] = m.z;
    const V3 a = induction(P, uu, aa);
    out[4] = a.x;
    out[5] = a.y;
    out[6] = a.z;
    out[7] = P.entropy ? entropy(P, ss, uu, lnrho, aa) : (Scalar)0;
}

/* ref src/utils/memory.cc:75-149: halo cell <- wrapped interior cell. */
template <typename T>
static void periodic_fill(T* const* f, int nfields, const Grid& g)
{
    for (int w = 0; w < nfields; ++w) {
        T* a = f[w];
        for (int k = 0; k < g.mz; ++k)
            for (int j = 0; j < g.my; ++j)
                for (int i = 0; i < g.mx; ++i) {
                    if (i >= 3 && i < g.nx + 3 && j >= 3 && j < g.ny + 3 && k >= 3 && k < g.nz + 3)
                        continue;
                    const int is = ((i - 3) + g.nx) % g.nx + 3;
                    const int js = ((j - 3) + g.ny) % g.ny + 3;
                    const int ks = ((k - 3) + g.nz) % g.nz + 3;
                    a[g.idx(i, j, k)] = a[g.idx(is, js, ks)];
                }
    }
}

/* ref modelsolver.cc:1214-1272 */
template <typename T>
static int host_integrate_step(T* const* f, const OracleConfig* cfg, double dt_in)
{
    const Grid g = make_grid(cfg);
    const Phys P = make_phys(cfg);
    const T dt = (T)dt_in;
    const long count = (long)g.mx * g.my * g.mz;
    std::vector<T> wstore((size_t)count * 8, (T)0);
    T* w[8];
    for (int q = 0; q < 8; ++q)
        w[q] = wstore.data() + (size_t)q * count;
    const Scalar alpha[3] = {(Scalar)0, (Scalar)-5 / (Scalar)9, (Scalar)-153 / (Scalar)128};
    const Scalar beta[3] = {(Scalar)1 / (Scalar)3, (Scalar)15 / (Scalar)16, (Scalar)8 / (Scalar)15};
    for (int s = 0; s < 3; ++s) {
        periodic_fill(f, 8, g);
#pragma omp parallel for schedule(static)
        for (int k = 3; k < g.nz + 3; ++k)
            for (int j = 3; j < g.ny + 3; ++j)
                for (int i = 3; i < g.nx + 3; ++i) {
                    Scalar r[8];
                    rhs_at(f, g, P, i, j, k, r);
                    const long c = g.idx(i, j, k);
                    for (int q = 0; q < 8; ++q) {
                        if (s == 0)
                            w[q][c] = (T)(r[q] * (Scalar)dt);
                        else
                            w[q][c] = (T)(alpha[s] * (Scalar)w[q][c] + r[q] * (Scalar)dt);
                    }
                }
#pragma omp parallel for schedule(static)
        for (int k = 3; k < g.nz + 3; ++k)
            for (int j = 3; j < g.ny + 3; ++j)
                for (int i = 3; i < g.nx + 3; ++i) {
                    const long c = g.idx(i, j, k);
                    for (int q = 0; q < 8; ++q)
                        f[q][c] += (T)(beta[s] * (Scalar)w[q][c]);
                }
    }
    return 0;
}

template <typename T>
static void rhs_field(T* const* f, const OracleConfig* cfg, T* const* out)
{
    const Grid g = make_grid(cfg);
    const Phys P = make_phys(cfg);
#pragma omp parallel for schedule(static)
    for (int k = 3; k < g.nz + 3; ++k)
        for (int j = 3; j < g.ny + 3; ++j)
            for (int i = 3; i < g.nx + 3; ++i) {
                Scalar r[8];
                rhs_at(f, g, P, i, j, k, r);
                for (int q = 0; q < 8; ++q)
                    out[q][g.idx(i, j, k)] = (T)r[q];
            }
}

/* ---------------- reductions (ref modelreduce.cc:108-340) ---------------- */
static inline long double map_value(int map, const long double* v)
{
    switch (map) {
    case 1: return v[0];                                            /* AC_MAP_VTXBUF */
    case 2: return v[0] * v[0];                                     /* _SQUARE */
    case 3: return expl(v[0]) * expl(v[0]);                         /* _EXP_SQUARE */
    case 4: return sqrtl(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);  /* VTXBUF3_NORM */
    case 5: return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];         /* VTXBUF3_SQUARE */
    case 6: return expl(v[0]) * expl(v[0]) + expl(v[1]) * expl(v[1]) + expl(v[2]) * expl(v[2]);
    case 7: return sqrtl(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / sqrtl(expl(v[3])); /* ALFVEN */
    case 8: return (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / expl(v[3]);
    default: return 0;
    }
}

template <typename T>
static long double host_reduce(T* const* bufs, int nbufs, int map, int op, int rms,
                               const OracleConfig* cfg)
{
    const Grid g = make_grid(cfg);
    long double v[4] = {0, 0, 0, 0};
    const long c0 = g.idx(3, 3, 3);
    for (int q = 0; q < nbufs; ++q)
        v[q] = (long double)bufs[q][c0];
    long double res = (op == 3) ? 0.0L : map_value(map, v);
    for (int k = 3; k < g.nz + 3; ++k)
        for (int j = 3; j < g.ny + 3; ++j)
            for (int i = 3; i < g.nx + 3; ++i) {
                const long c = g.idx(i, j, k);
                for (int q = 0; q < nbufs; ++q)
                    v[q] = (long double)bufs[q][c];
                const long double x = map_value(map, v);
                if (op == 1)
                    res = x < res ? x : res;
                else if (op == 2)
                    res = x > res ? x : res;
                else
                    res = res + x;
            }
    if (rms) {
        const long double inv_n = 1.0L / ((long double)g.nx * g.ny * g.nz);
        return sqrtl(inv_n * res);
    }
    return res;
}

/* ---------------- ref-cpu: the reference CPU-build algorithm in FP64 ----------------
 * Two passes per substep like twopass_solve_intermediate / twopass_solve_final
 * (mhdsolver.ac:586-636): pass 1 computes w = alpha*w + dt*rhs from `in` stencils, pass 2
 * computes f = f + beta*w. Stencil coefficients are folded with inv_ds first, as the DSL's
 * Stencil declarations do (derivs.h:591-752), and evaluated in double. */
template <typename T>
struct DPhys {
    T cs2_sound, nu, zeta, eta, mu0, inv_mu0, cp, inv_cp, gamma, lnT0, lnrho0, K;
    bool upwd, entropy;
};

template <typename T>
struct DStencils {
    T d1[4][3];   /* per axis, per offset 1..3 (antisymmetric) */
    T d2[4][3];   /* [0] centre, per axis */
    T d2c[3];
    T up[4][3];   /* [0] centre */
    T upc[3];
    T mx[4][3];   /* per pair (xy, xz, yz), per offset 1..3 */
};

template <int AX, typename T>
static inline T dsten1(const T* a, long c, const long* s, const DStencils<T>& S)
{
    const long st = s[AX];
    return S.d1[1][AX] * (a[c + st] - a[c - st]) + S.d1[2][AX] * (a[c + 2 * st] - a[c - 2 * st]) +
           S.d1[3][AX] * (a[c + 3 * st] - a[c - 3 * st]);
}

template <typename T>
static void refcpu_substep(T* const* in, T* const* wbuf, const Grid& g,
                           const DPhys<T>& P, const DStencils<T>& S, int step, T dt, int pass)
{
    const T alpha[3] = {T(0.0), T(-5.0 / 9.0), T(-153.0 / 128.0)};
    const T beta[3] = {T(1.0 / 3.0), T(15.0 / 16.0), T(8.0 / 15.0)};
    const long s[3] = {1, g.sy, g.sz};
#pragma omp parallel for schedule(static)
    for (int k = 3; k < g.nz + 3; ++k)
        for (int j = 3; j < g.ny + 3; ++j)
            for (int i = 3; i < g.nx + 3; ++i) {
                const long c = g.idx(i, j, k);
                if (pass == 1) {
                    for (int q = 0; q < 8; ++q)
                        in[q][c] = in[q][c] + beta[step] * wbuf[q][c];
                    continue;
                }
                T val[8], gr[8][3], lap[8], upw[8][3];
                for (int q = 0; q < 8; ++q) {
                    const T* a = in[q];
                    val[q] = a[c];
                    gr[q][0] = dsten1<0, T>(a, c, s, S);
                    gr[q][1] = dsten1<1, T>(a, c, s, S);
                    gr[q][2] = dsten1<2, T>(a, c, s, S);
                    T l = T(0.0);
                    for (int ax = 0; ax < 3; ++ax) {
                        const long st = s[ax];
                        const T s1 = a[c + st] + a[c - st], s2 = a[c + 2 * st] + a[c - 2 * st],
                                     s3 = a[c + 3 * st] + a[c - 3 * st];
                        l += S.d2c[ax] * a[c] + S.d2[1][ax] * s1 + S.d2[2][ax] * s2 +
                             S.d2[3][ax] * s3;
                        upw[q][ax] = S.upc[ax] * a[c] + S.up[1][ax] * s1 + S.up[2][ax] * s2 +
                                     S.up[3][ax] * s3;
                    }
                    lap[q] = l;
                }
                auto mixed = [&](const T* a, int p, int ax, int bx) {
                    const long sa = s[ax], sb = s[bx];
                    T r = T(0.0);
                    for (int o = 1; o <= 3; ++o)
                        r += S.mx[o][p] * (a[c + o * sa + o * sb] + a[c - o * sa - o * sb] -
                                           a[c + o * sa - o * sb] - a[c - o * sa + o * sb]);
                    return r;
                };
                auto second = [&](const T* a, int ax) {
                    const long st = s[ax];
                    return S.d2c[ax] * a[c] + S.d2[1][ax] * (a[c + st] + a[c - st]) +
                           S.d2[2][ax] * (a[c + 2 * st] + a[c - 2 * st]) +
                           S.d2[3][ax] * (a[c + 3 * st] + a[c - 3 * st]);
                };
                /* grad div of a vector stored in fields b, b+1, b+2 */
                auto graddiv = [&](int b, T gd[3]) {
                    gd[0] = second(in[b], 0) + mixed(in[b + 1], 0, 0, 1) + mixed(in[b + 2], 1, 0, 2);
                    gd[1] = mixed(in[b], 0, 0, 1) + second(in[b + 1], 1) + mixed(in[b + 2], 2, 1, 2);
                    gd[2] = mixed(in[b], 1, 0, 2) + mixed(in[b + 1], 2, 1, 2) + second(in[b + 2], 2);
                };
                T gdu[3], gda[3];
                graddiv(1, gdu);
                graddiv(4, gda);
                const T ux = val[1], uy = val[2], uz = val[3];
                const T au[3] = {std::fabs(ux), std::fabs(uy), std::fabs(uz)};
                const T u[3] = {ux, uy, uz};
                const T divu = gr[1][0] + gr[2][1] + gr[3][2];
                T rhs[8];
                rhs[0] = -(u[0] * gr[0][0] + u[1] * gr[0][1] + u[2] * gr[0][2]) - divu;
                if (P.upwd)
                    rhs[0] += au[0] * upw[0][0] + au[1] * upw[0][1] + au[2] * upw[0][2];
                T Sm[3][3];
                Sm[0][0] = T(2.0 / 3.0) * gr[1][0] - T(1.0 / 3.0) * (gr[2][1] + gr[3][2]);
                Sm[0][1] = T(0.5) * (gr[1][1] + gr[2][0]);
                Sm[0][2] = T(0.5) * (gr[1][2] + gr[3][0]);
                Sm[1][1] = T(2.0 / 3.0) * gr[2][1] - T(1.0 / 3.0) * (gr[1][0] + gr[3][2]);
                Sm[1][2] = T(0.5) * (gr[2][2] + gr[3][1]);
                Sm[2][2] = T(2.0 / 3.0) * gr[3][2] - T(1.0 / 3.0) * (gr[1][0] + gr[2][1]);
                Sm[1][0] = Sm[0][1];
                Sm[2][0] = Sm[0][2];
                Sm[2][1] = Sm[1][2];
                const T cs2 = P.entropy ? P.cs2_sound * std::exp(P.gamma * val[7] / P.cp +
                                                                 (P.gamma - T(1.0)) * (val[0] - P.lnrho0))
                                             : P.cs2_sound;
                const T jv[3] = {P.inv_mu0 * (gda[0] - lap[4]), P.inv_mu0 * (gda[1] - lap[5]),
                                      P.inv_mu0 * (gda[2] - lap[6])};
                const T B[3] = {gr[6][1] - gr[5][2], gr[4][2] - gr[6][0], gr[5][0] - gr[4][1]};
                const T inv_rho = T(1.0) / std::exp(val[0]);
                const T jxB[3] = {jv[1] * B[2] - jv[2] * B[1], jv[2] * B[0] - jv[0] * B[2],
                                       jv[0] * B[1] - jv[1] * B[0]};
                for (int d = 0; d < 3; ++d) {
                    const int q = 1 + d;
                    T m = -(gr[q][0] * u[0] + gr[q][1] * u[1] + gr[q][2] * u[2]);
                    if (P.upwd)
                        m += upw[q][0] * au[0] + upw[q][1] * au[1] + upw[q][2] * au[2];
                    const T press = P.entropy ? P.inv_cp * gr[7][d] + gr[0][d] : gr[0][d];
                    m -= cs2 * press;
                    m += inv_rho * jxB[d];
                    m += P.nu * (lap[q] + T(1.0 / 3.0) * gdu[d] +
                                 T(2.0) * (Sm[d][0] * gr[0][0] + Sm[d][1] * gr[0][1] + Sm[d][2] * gr[0][2]));
                    m += P.zeta * gdu[d];
                    rhs[q] = m;
                }
                rhs[4] = (u[1] * B[2] - u[2] * B[1]) + P.eta * lap[4];
                rhs[5] = (u[2] * B[0] - u[0] * B[2]) + P.eta * lap[5];
                rhs[6] = (u[0] * B[1] - u[1] * B[0]) + P.eta * lap[6];
                if (P.entropy) {
                    const T lnT = P.lnT0 + P.gamma * val[7] / P.cp + (P.gamma - T(1.0)) * (val[0] - P.lnrho0);
                    const T inv_pT = T(1.0) / (std::exp(val[0]) * std::exp(lnT));
                    T cS = T(0.0);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            cS += Sm[a][b] * Sm[a][b];
                    const T R = P.eta * P.mu0 * (jv[0] * jv[0] + jv[1] * jv[1] + jv[2] * jv[2]) +
                                     T(2.0) * std::exp(val[0]) * P.nu * cS + P.zeta * std::exp(val[0]) * divu * divu;
                    const T first = P.gamma * P.inv_cp * lap[7] + (P.gamma - T(1.0)) * lap[0];
                    T dt23 = T(0.0);
                    for (int d = 0; d < 3; ++d) {
                        const T sec = P.gamma * P.inv_cp * gr[7][d] + (P.gamma - T(1.0)) * gr[0][d];
                        const T thi = P.gamma * (P.inv_cp * gr[7][d] + gr[0][d]) - gr[0][d];
                        dt23 += sec * thi;
                    }
                    const T chi = P.K / (std::exp(val[0]) * P.cp);
                    rhs[7] = -(u[0] * gr[7][0] + u[1] * gr[7][1] + u[2] * gr[7][2]) + inv_pT * R +
                             P.cp * chi * (first + dt23);
                }
                else {
                    rhs[7] = T(0.0);
                }
                for (int q = 0; q < 8; ++q)
                    wbuf[q][c] = alpha[step] * wbuf[q][c] + rhs[q] * dt;
            }
}

} // namespace

/* One full RK3 step of the ref-cpu algorithm in arithmetic type T (the reference's CPU build with
 * AcReal = T: two passes per substep, OpenMP over z with nthreads threads; nthreads <= 0 keeps
 * the OpenMP default). `wbuf` is the caller's 8-field intermediate buffer of the same shape. */
template <typename T>
static int refcpu_integrate_step(T* const* fields, T* const* wbuf, const OracleConfig* cfg,
                                 double dt, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const Grid g = make_grid(cfg);
    DPhys<T> P;
    P.cs2_sound = T(cfg->cs_sound * cfg->cs_sound);
    P.nu = T(cfg->nu_visc);
    P.zeta = T(cfg->zeta);
    P.eta = T(cfg->eta);
    P.mu0 = T(cfg->mu0);
    P.inv_mu0 = T(1.0 / cfg->mu0);
    P.cp = T(cfg->cp_sound);
    P.inv_cp = T(1.0 / cfg->cp_sound);
    P.gamma = T(cfg->gamma);
    P.lnT0 = T(cfg->lnT0);
    P.lnrho0 = T(cfg->lnrho0);
    P.K = T(cfg->K_heatcond);
    P.upwd = cfg->lupwd != 0;
    P.entropy = cfg->lentropy != 0;
    DStencils<T> S;
    const double c1[4] = {0.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
    const double c2[4] = {-49.0 / 18.0, 3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0};
    const double cu[4] = {-20.0 / 60.0, 15.0 / 60.0, -6.0 / 60.0, 1.0 / 60.0};
    const double cm[4] = {0.0, 270.0 / 720.0, -27.0 / 720.0, 2.0 / 720.0};
    double inv[3];
    for (int d = 0; d < 3; ++d)
        inv[d] = 1.0 / cfg->ds[d];
    for (int d = 0; d < 3; ++d) {
        for (int o = 1; o <= 3; ++o) {
            S.d1[o][d] = T(inv[d] * c1[o]);
            S.d2[o][d] = T(inv[d] * inv[d] * c2[o]);
            S.up[o][d] = T(inv[d] * cu[o]);
        }
        S.d2c[d] = T(inv[d] * inv[d] * c2[0]);
        S.upc[d] = T(inv[d] * cu[0]);
    }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int p = 0; p < 3; ++p)
        for (int o = 1; o <= 3; ++o)
            S.mx[o][p] = T(inv[pairs[p][0]] * inv[pairs[p][1]] * cm[o]);
    for (int s = 0; s < 3; ++s) {
        periodic_fill(fields, 8, g);
        refcpu_substep<T>(fields, wbuf, g, P, S, s, T(dt), 0);
        refcpu_substep<T>(fields, wbuf, g, P, S, s, T(dt), 1);
    }
    return 0;
}

extern "C" {

int oracle_integrate_step_f64(double* const* fields, const OracleConfig* cfg, double dt)
{
    return host_integrate_step<double>(fields, cfg, dt);
}
int oracle_integrate_step_f32(float* const* fields, const OracleConfig* cfg, double dt)
{
    return host_integrate_step<float>(fields, cfg, dt);
}
void oracle_periodic_fill_f64(double* const* fields, int nfields, const OracleConfig* cfg)
{
    periodic_fill<double>(fields, nfields, make_grid(cfg));
}
void oracle_periodic_fill_f32(float* const* fields, int nfields, const OracleConfig* cfg)
{
    periodic_fill<float>(fields, nfields, make_grid(cfg));
}
/* RHS of every field at every interior cell, long double arithmetic rounded to double.
 * Halos of `fields` must be filled by the caller. */
void oracle_rhs_f64(double* const* fields, const OracleConfig* cfg, double* const* out)
{
    rhs_field<double>(fields, cfg, out);
}
long double oracle_reduce_f64(double* const* bufs, int nbufs, int map, int op, int rms,
                              const OracleConfig* cfg)
{
    return host_reduce<double>(bufs, nbufs, map, op, rms, cfg);
}
long double oracle_reduce_f32(float* const* bufs, int nbufs, int map, int op, int rms,
                              const OracleConfig* cfg)
{
    return host_reduce<float>(bufs, nbufs, map, op, rms, cfg);
}

int refcpu_integrate_step_f64(double* const* fields, double* const* wbuf, const OracleConfig* cfg,
                              double dt, int nthreads)
{
    return refcpu_integrate_step<double>(fields, wbuf, cfg, dt, nthreads);
}

int refcpu_integrate_step_f32(float* const* fields, float* const* wbuf, const OracleConfig* cfg,
                              double dt, int nthreads)
{
    return refcpu_integrate_step<float>(fields, wbuf, cfg, dt, nthreads);
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

} // extern "C"
