// solve_sweep.hip -- the fused RK3 substep as an unrolled z-sweep (FP32, gfx950).
//
// Same semantics as solve_direct.hip (ref acc-runtime/samples/mhd_modular/mhdsolver.ac:586-636,
// one fused pass) and the same tiling as the delay-line kernel of solve_tiled.hip:
//
//  * A workgroup owns a TX x TY column block of the interior and a chunk of ZC planes and
//    marches upwards in z. Each plane step DMAs one xy plane of all eight fields, with its
//    radius-3 halo, straight into LDS (global_load_lds, no staging registers).
//  * Every thread owns one (x, y) column: its in-plane stencils come from the LDS plane; what
//    needs other planes stays in registers -- a 7-deep queue of the column's values (z first,
//    second and upwind derivatives), push-accumulated chains for the xz / yz mixed derivatives
//        d_xz f(z) = sum_i c_i [ Dx_i f(z+i) - Dx_i f(z-i) ],  Dx_i f = f(x+i) - f(x-i),
//    of grad(div u) and grad(div A), and a 3-plane delay line of the in-plane derivatives (the
//    output plane z lags the newest plane p = z + 3). The A / ss entries of the delay line ride
//    a per-thread ring in LDS.
//
// What differs from the delay-line kernel is how the plane loop maps to the CDNA4 issue model:
//
//  * Two plane steps are unrolled into one loop iteration, and the two LDS plane buffers are two
//    separate arrays used in static alternation. The compiler then knows that a plane's stencil
//    reads do not alias the DMA that fills the other buffer and issues them at once; with one
//    array indexed by a run-time buffer number it waits for the whole DMA (vmcnt(0)) before the
//    first stencil read of every plane step, exposing its full latency.
//  * Inside an iteration every queue, chain and delay entry is named statically; the entries are
//    shifted once per two planes instead of once per plane.
//  * w (and the compensation record) of the next plane step's output cell is loaded right after
//    this step's output, so its latency hides behind the in-plane stage and the barrier that
//    waits for the DMA anyway (a load issued after the DMA and consumed before the barrier would
//    make the compiler wait for the DMA too).
//  * ISO: on a grid with equal spacing along x, y and z (every configuration of the benchmark
//    suite) one set of stencil coefficients serves all three axes, which keeps the kernel within
//    the scalar register file.
//  * The carried state is pinned to registers at the end of each plane step (empty asm), so the
//    optimiser cannot move computations across plane steps and lengthen their live ranges; the
//    kernel fits the 256 registers of two waves per SIMD without spills.
//
// No MFMA: a stencil is not a dense contraction. Workgroups are numbered so that consecutive
// tiles (which share halo rows) land on the same XCD and its L2: workgroup b runs on XCD b % 8.
#include <type_traits>

#include "kernels.h"

namespace acamd {

namespace {

// In-plane quantities of one plane that its output needs three plane steps later (lnrho and u;
// the ten of A and ss ride the LDS ring).
template <typename T>
struct SweepPlane {
    T lr_dx, lr_dy, lr_lap, lr_up;
    T u_dx[3], u_dy[3], u_lap[3], u_up[3];
};

constexpr int SW_NDL = 10;  // delayed A / ss values in the LDS ring
constexpr int SW_NCH = 6;   // chains: grad(div u) x, y, z ; grad(div A) x, y, z

}  // namespace

template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY, bool WRAP, int STEP, bool ISO>
__device__ __forceinline__ void sweep_body(const SolveArgs<T>& a, int zchunk, int ntx, int nty,
                                           int nchunks)
{
    static_assert(sizeof(T) == 4, "the sweep kernel is the FP32 path");
    constexpr int UNR = 2;  // plane steps per loop iteration (the LDS buffers alternate statically)
    constexpr int LX = TX + 2 * R;
    constexpr int LY = TY + 2 * R;
    constexpr int LP = LX * LY;                 // one field-plane with halo (elements = dwords)
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;                 // waves per workgroup
    constexpr int NCHUNK = (LP + 63) / 64;      // wave-wide DMA chunks of one field-plane
    constexpr int NR = (NCHUNK + NW - 1) / NW;  // chunks per wave per field-plane
    // Each field-plane occupies NR * NW whole chunks, so every wave issues exactly NR DMA
    // instructions per field with all lanes active: lanes past the plane's last element read a
    // valid address and land in the padding.
    constexpr int LPS = NR * NW * 64;
    // 32-bit compensation: second, upwind and mixed derivatives from differences to the centre
    // value (exact for neighbouring values) rather than sums of neighbours.
    constexpr bool CANCEL_FREE = KB == 32;
    // The first and last three planes of a chunk are not output: they only feed the z-stencils
    // and chains of their neighbours, so their in-plane stage is skipped.
    constexpr bool SKIP_EDGE = true;
    static_assert(NT % 64 == 0, "workgroup must be whole waves");
    __shared__ T lds0[NF * LPS];
    __shared__ T lds1[NF * LPS];
    __shared__ T dl[3][SW_NDL][NT];  // delayed A / ss values, [slot][value][thread]

    // ---- XCD-aware tile numbering: workgroup b runs on XCD b % 8; each XCD gets a contiguous
    // range of tiles (x fastest, then y, then z-chunk), so halo rows are shared within one L2.
    const int nblocks = ntx * nty * nchunks;
    const int bid = blockIdx.x;
    int logical = bid;
    if ((nblocks & 7) == 0)
        logical = (bid & 7) * (nblocks >> 3) + (bid >> 3);
    const int tile_x = logical % ntx;
    const int rest = logical / ntx;
    const int tile_y = rest % nty;
    const int chunk = rest / nty;

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + TX * ty;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: M0 is scalar math
    const int lane = tid & 63;
    const int x0 = tile_x * TX, y0 = tile_y * TY, z0 = chunk * zchunk;
    const long long sz = a.sz;

    // ---- DMA slots of this lane: chunk j = wave + r * NW covers the elements [64 j, 64 j + 64)
    // of a field-plane; lane l moves element e = 64 j + l.
    unsigned g_off[NR];  // byte offset of that element within the array plane
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        int e = (wave + r * NW) * 64 + lane;
        if (e >= LP)
            e = 0;  // padding lane: re-read the first element (in bounds, value unused)
        const int ly = e / LX, lx = e - ly * LX;
        int hx = x0 + lx, hy = y0 + ly;  // array coordinates of the element
        if (WRAP) {
            // halo elements from their periodic images in the interior
            hx += hx < R ? a.nx : (hx >= a.nx + R ? -a.nx : 0);
            hy += hy < R ? a.ny : (hy >= a.ny + R ? -a.ny : 0);
        }
        g_off[r] = (unsigned)((hx + a.mx * hy) * 4);
    }
    const char* const in0 = (const char*)a.in[0];
    char* const out0 = (char*)a.out[0];
    char* const w0 = (char*)a.w[0];
    char* const comp0 = (char*)a.comp[0];
    // With 16- and 32-bit compensation the per-field bases of the RK arrays are pinned to scalar
    // registers, so loads and stores take the uniform base + 32-bit lane offset form.
    constexpr bool SBASE = KB >= 16;
    auto field_base = [](char* p) {
        if constexpr (SBASE)
            return sbase(p);
        else
            return (T*)p;
    };
    auto ld_field = [](char* p, unsigned boff) -> T {
        if constexpr (SBASE)
            return ld_g<T>(sbase(p), boff);
        else
            return ld_b((const T*)p, boff);
    };
    // DMA of interior plane p (array z index p + R) into the LDS buffer dst.
    auto dma_plane = [&](T* dst, int p, long long fsb) __attribute__((always_inline)) {
        if (WRAP)
            p += p < 0 ? a.nz : (p >= a.nz ? -a.nz : 0);
        const long long base = sz * (long long)(p + R) * 4;
        unsigned go[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r)
            go[r] = opaque(g_off[r]);  // one copy per plane, not one 64-bit address per field
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const char* src = in0 + (base + f * fsb);
            char* d = (char*)&dst[f * LPS];
#pragma unroll
            for (int r = 0; r < NR; ++r)
                __builtin_amdgcn_global_load_lds((const void*)(src + go[r]),
                                                 (lds_ptr_t)(d + 256 * (wave + r * NW)), 4, 0, 0);
        }
    };

    const Coeffs<T>& C = a.C;
    // stencil coefficients of axis ax (x, y, z) and of the mixed plane pl (xy, xz, yz)
    constexpr int AY = ISO ? 0 : 1, AZ = ISO ? 0 : 2;
    // Register state, indexed by plane relative to the first plane step of the current
    // unrolled iteration (whose newest plane is p0):
    //   q[f][k]   the column's value of field f at plane p0 - 6 + k, k in [0, 6 + UNR)
    //   ch[c][k]  chain c of plane p0 - 3 + k, k in [0, 6 + UNR)
    //   dly[k]    delayed in-plane quantities of plane p0 - 3 + k, k in [0, 3 + UNR)
    // Plane step u (newest plane p0 + u) reads q[.][u .. u + 6], outputs plane p0 + u - 3 from
    // ch[.][u] and dly[u], pushes into ch[.][u .. u + 6] and writes dly[u + 3]. At the end of the
    // iteration the entries still needed move down by UNR.
    T q[NF][6 + UNR];
    T ch[SW_NCH][6 + UNR];
    SweepPlane<T> dly[3 + UNR] = {};  // zeroed: no carried register is read before written
    T wq[2][NF];            // w of the output cell, loaded one plane step ahead
    CompRec<KB> cq[2];      // its compensation record (KB > 0)
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < 6; ++k)
            q[f][k] = T(0);
#pragma unroll
    for (int c = 0; c < SW_NCH; ++c)
#pragma unroll
        for (int k = 0; k < 6; ++k)
            ch[c][k] = T(0);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        wq[0][f] = T(0);
        wq[1][f] = T(0);
    }
    cq[0] = CompRec<KB>{};
    cq[1] = CompRec<KB>{};

    const int step = STEP >= 0 ? STEP : a.step;
    const T dt = solve_dt(a);
    double u2max = 0.0;  // max |u_new|^2 (the courant timestep mode, substep 2)
    const int gx = x0 + tx, gy = y0 + ty;
    const int ctr = (ty + R) * LX + (tx + R);
    unsigned ioff = (unsigned)(((gx + R) + a.mx * (gy + R)) * 4);  // in-plane byte offset
    int dslot = 0;  // LDS ring slot of the newest plane (and of the output plane three steps on)
    const int niter = zchunk + 2 * R;

    dma_plane(lds0, z0 - R, a.fstride * 4);
    plane_barrier();

    for (int it0 = 0;; it0 += UNR) {
        auto plane_step = [&](auto UC) __attribute__((always_inline)) {
            constexpr int u = decltype(UC)::value;
            T* const lcur = (u & 1) ? lds1 : lds0;  // holds plane p
            T* const lnxt = (u & 1) ? lds0 : lds1;  // receives plane p + 1
            const int it = it0 + u;
            const int p = z0 - R + it;
            ioff = opaque(ioff);
            const long long fsb = opaque_s64(a.fstride * 4);  // field stride, bytes

            // ---- (0) DMA of the next plane
            if (it + 1 < niter)
                dma_plane(lnxt, p + 1, fsb);
            // this thread's centre in the LDS planes, recomputed per plane step (held across the
            // loop, the addresses of both buffers would take registers the state needs)
            const unsigned lctr = opaque((unsigned)ctr * 4u);
            auto LD = [&](int f, int dx, int dy) __attribute__((always_inline)) -> T {
                return *(const T*)((const char*)lcur + lctr + 4 * (f * LPS + dy * LX + dx));
            };
#pragma unroll
            for (int f = 0; f < NF; ++f)
                q[f][u + 6] = LD(f, 0, 0);
            SweepPlane<T>& nv = dly[u + 3];
            T az_dx_p = T(0), az_dy_p = T(0);  // of plane p, stored to the LDS ring below

            // ---- (1) z-chains: plane p adds +c_i D(p) to plane p - i and -c_i D(p) to p + i
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int fx = 1 + 3 * v, fy = 2 + 3 * v, fz = 3 + 3 * v;
                T dxz = T(0), dyz = T(0), dxx_ = T(0), dyy_ = T(0);
#pragma unroll
                for (int i = 1; i <= 3; ++i) {
                    const T Dxz = LD(fz, i, 0) - LD(fz, -i, 0);
                    const T Dyz = LD(fz, 0, i) - LD(fz, 0, -i);
                    const T Dxx = LD(fx, i, 0) - LD(fx, -i, 0);
                    const T Dyy = LD(fy, 0, i) - LD(fy, 0, -i);
                    const T kx = C.mx[AY][i], ky = C.mx[AZ][i];
                    const T ez = kx * Dxx + ky * Dyy;  // d_xz v_x + d_yz v_y -> (grad div v)_z
                    // (grad div v)_x gets d_xz v_z, (grad div v)_y gets d_yz v_z
                    ch[3 * v + 0][u + 3 - i] = __builtin_fmaf(kx, Dxz, ch[3 * v + 0][u + 3 - i]);
                    ch[3 * v + 1][u + 3 - i] = __builtin_fmaf(ky, Dyz, ch[3 * v + 1][u + 3 - i]);
                    ch[3 * v + 2][u + 3 - i] += ez;
                    if (i == 3) {  // the first contribution to plane p + 3
                        ch[3 * v + 0][u + 6] = -(kx * Dxz);
                        ch[3 * v + 1][u + 6] = -(ky * Dyz);
                        ch[3 * v + 2][u + 6] = -ez;
                    }
                    else {
                        ch[3 * v + 0][u + 3 + i] = __builtin_fmaf(-kx, Dxz, ch[3 * v + 0][u + 3 + i]);
                        ch[3 * v + 1][u + 3 + i] = __builtin_fmaf(-ky, Dyz, ch[3 * v + 1][u + 3 + i]);
                        ch[3 * v + 2][u + 3 + i] -= ez;
                    }
                    dxz += C.d1[0][i] * Dxz;
                    dyz += C.d1[AY][i] * Dyz;
                    dxx_ += C.d1[0][i] * Dxx;
                    dyy_ += C.d1[AY][i] * Dyy;
                }
                if (v == 0) {
                    nv.u_dx[2] = dxz;
                    nv.u_dy[2] = dyz;
                    nv.u_dx[0] = dxx_;
                    nv.u_dy[1] = dyy_;
                }
                else {
                    az_dx_p = dxz;
                    az_dy_p = dyz;
                }
            }

            // ---- (2) output plane z = p - 3
            if (it >= 2 * R) {
                const SweepPlane<T>& o = dly[u];
                T dv[SW_NDL];  // the A / ss values of plane z, written three steps ago
#pragma unroll
                for (int k = 0; k < SW_NDL; ++k)
                    dv[k] = dl[dslot][k][tid];
                T dz[NF], dzz[NF], upz[4];
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const T c0 = q[f][u + 3];
                    T D[4], S[4];
#pragma unroll
                    for (int i = 1; i <= 3; ++i) {
                        const T zp = q[f][u + 3 + i], zm = q[f][u + 3 - i];
                        D[i] = zp - zm;
                        S[i] = CANCEL_FREE ? (zp - c0) + (zm - c0) : zp + zm;
                    }
                    dz[f] = C.d1[AZ][1] * D[1] + C.d1[AZ][2] * D[2] + C.d1[AZ][3] * D[3];
                    if constexpr (CANCEL_FREE) {
                        dzz[f] = C.d2[AZ][1] * S[1] + C.d2[AZ][2] * S[2] + C.d2[AZ][3] * S[3];
                        if (UPWD && f < 4)
                            upz[f] = C.up[AZ][1] * S[1] + C.up[AZ][2] * S[2] + C.up[AZ][3] * S[3];
                    }
                    else {
                        dzz[f] = C.d2[AZ][0] * c0 + C.d2[AZ][1] * S[1] + C.d2[AZ][2] * S[2] +
                                 C.d2[AZ][3] * S[3];
                        if (UPWD && f < 4)
                            upz[f] = C.up[AZ][0] * c0 + C.up[AZ][1] * S[1] + C.up[AZ][2] * S[2] +
                                     C.up[AZ][3] * S[3];
                    }
                }
                CellDerivs<T> d;
                d.lr = q[0][u + 3];
                d.glr[0] = o.lr_dx;
                d.glr[1] = o.lr_dy;
                d.glr[2] = dz[0];
                d.laplr = o.lr_lap + dzz[0];
                const T auz = ac_abs(q[3][u + 3]);
                d.upl = UPWD ? o.lr_up + auz * upz[0] : T(0);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    d.u[i] = q[1 + i][u + 3];
                    d.gu[i][0] = o.u_dx[i];
                    d.gu[i][1] = o.u_dy[i];
                    d.gu[i][2] = dz[1 + i];
                    d.lapu[i] = o.u_lap[i] + dzz[1 + i];
                    d.upu[i] = UPWD ? o.u_up[i] + auz * upz[1 + i] : T(0);
                    d.lapa[i] = dv[4 + i] + dzz[4 + i];
                }
                d.gdu[0] = ch[0][u];
                d.gdu[1] = ch[1][u];
                d.gdu[2] = ch[2][u] + dzz[3];
                d.gda[0] = ch[3][u];
                d.gda[1] = ch[4][u];
                d.gda[2] = ch[5][u] + dzz[6];
                d.ga[0][0] = T(0);
                d.ga[0][1] = dv[0];
                d.ga[0][2] = dz[4];
                d.ga[1][0] = dv[1];
                d.ga[1][1] = T(0);
                d.ga[1][2] = dz[5];
                d.ga[2][0] = dv[2];
                d.ga[2][1] = dv[3];
                d.ga[2][2] = T(0);
                d.s = q[7][u + 3];
                d.gs[0] = dv[7];
                d.gs[1] = dv[8];
                d.gs[2] = dz[7];
                d.laps = dv[9] + dzz[7];

                T rhs[NF];
                mhd_rhs<T, UPWD, ENTROPY, CANCEL_FREE>(d, a.P, rhs);
                const long long poff = sz * (long long)p;  // plane z's array index is p - R + R
                T fnew[NF];
                const CompRec<KB>& crec = cq[u & 1];
                CompRec<KB> cout = crec;  // ss keeps its record entry without entropy
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const long long fo = f * fsb + poff * 4;  // bytes
                    auto op = field_base(out0 + fo);
                    if (f == 7 && !ENTROPY) {
                        fnew[f] = q[f][u + 3];
                        st_b(op, ioff, q[f][u + 3]);
                        continue;
                    }
                    auto wp = field_base(w0 + fo);
                    int cenc = 0;
                    fnew[f] = rk_update_pre<T, KB>(wp, op, ioff, step, a.alpha, a.beta, q[f][u + 3],
                                                   rhs[f], dt, wq[u & 1][f],
                                                   comp_rec_get<KB>(crec, f), &cenc);
                    if (KB > 0)
                        comp_rec_set<KB>(cout, f, cenc);
                }
                if (KB > 0)
                    comp_rec_store<KB>(comp0 + poff * (long long)(NF * KB / 8),
                                       ioff / 4u * (unsigned)(NF * KB / 8), cout);
                if (step == 2) {
                    const double u2 = (double)fnew[1] * (double)fnew[1] +
                                      (double)fnew[2] * (double)fnew[2] +
                                      (double)fnew[3] * (double)fnew[3];
                    u2max = fmax(u2max, u2);
                }
            }

            // ---- w / compensation of the next step's output cell
            if (it + 1 >= 2 * R && it + 1 < niter) {
                const long long poff = sz * (long long)(p + 1 - R + R);  // output plane of it + 1
                if (step > 0) {
#pragma unroll
                    for (int f = 0; f < NF; ++f)
                        wq[(u + 1) & 1][f] = ld_field(w0 + (f * fsb + poff * 4), ioff);
                }
                if (KB > 0)
                    cq[(u + 1) & 1] = comp_rec_load<KB>(comp0 + poff * (long long)(NF * KB / 8),
                                                         ioff / 4u * (unsigned)(NF * KB / 8));
            }

            // ---- (3) remaining in-plane derivatives of plane p, where plane p is itself output
            if (!SKIP_EDGE || (it >= R && it < zchunk + R)) {
                const T aux = ac_abs(q[1][u + 6]), auy = ac_abs(q[2][u + 6]);
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const T c0 = q[f][u + 6];
                    T Dx[4], Dy[4], Sx[4], Sy[4];
#pragma unroll
                    for (int i = 1; i <= 3; ++i) {
                        const T xp = LD(f, i, 0), xm = LD(f, -i, 0);
                        const T yp = LD(f, 0, i), ym = LD(f, 0, -i);
                        Dx[i] = xp - xm;
                        Dy[i] = yp - ym;
                        if constexpr (CANCEL_FREE) {
                            Sx[i] = (xp - c0) + (xm - c0);
                            Sy[i] = (yp - c0) + (ym - c0);
                        }
                        else {
                            Sx[i] = xp + xm;
                            Sy[i] = yp + ym;
                        }
                    }
                    T upxy = T(0);
                    T dxx, dyy;
                    if constexpr (CANCEL_FREE) {
                        dxx = C.d2[0][1] * Sx[1] + C.d2[0][2] * Sx[2] + C.d2[0][3] * Sx[3];
                        dyy = C.d2[AY][1] * Sy[1] + C.d2[AY][2] * Sy[2] + C.d2[AY][3] * Sy[3];
                        if (UPWD && f < 4) {
                            const T upx = C.up[0][1] * Sx[1] + C.up[0][2] * Sx[2] + C.up[0][3] * Sx[3];
                            const T upy = C.up[AY][1] * Sy[1] + C.up[AY][2] * Sy[2] + C.up[AY][3] * Sy[3];
                            upxy = aux * upx + auy * upy;
                        }
                    }
                    else {
                        dxx = C.d2[0][0] * c0 + C.d2[0][1] * Sx[1] + C.d2[0][2] * Sx[2] +
                              C.d2[0][3] * Sx[3];
                        dyy = C.d2[AY][0] * c0 + C.d2[AY][1] * Sy[1] + C.d2[AY][2] * Sy[2] +
                              C.d2[AY][3] * Sy[3];
                        if (UPWD && f < 4) {
                            const T upx = C.up[0][0] * c0 + C.up[0][1] * Sx[1] + C.up[0][2] * Sx[2] +
                                          C.up[0][3] * Sx[3];
                            const T upy = C.up[AY][0] * c0 + C.up[AY][1] * Sy[1] + C.up[AY][2] * Sy[2] +
                                          C.up[AY][3] * Sy[3];
                            upxy = aux * upx + auy * upy;
                        }
                    }
                    const bool need_dx = (f == 0 || f == 2 || f == 5 || f == 7);
                    const bool need_dy = (f == 0 || f == 1 || f == 4 || f == 7);
                    T dx = T(0), dy = T(0);
                    if (need_dx)
                        dx = C.d1[0][1] * Dx[1] + C.d1[0][2] * Dx[2] + C.d1[0][3] * Dx[3];
                    if (need_dy)
                        dy = C.d1[AY][1] * Dy[1] + C.d1[AY][2] * Dy[2] + C.d1[AY][3] * Dy[3];
                    if (f == 0) {
                        nv.lr_dx = dx;
                        nv.lr_dy = dy;
                        nv.lr_lap = dxx + dyy;
                        nv.lr_up = upxy;
                    }
                    else if (f < 4) {
                        if (f == 1)
                            nv.u_dy[0] = dy;
                        if (f == 2)
                            nv.u_dx[1] = dx;
                        nv.u_lap[f - 1] = dxx + dyy;
                        nv.u_up[f - 1] = upxy;
                    }
                    else if (f < 7) {
                        if (f == 4)
                            dl[dslot][0][tid] = dy;
                        if (f == 5)
                            dl[dslot][1][tid] = dx;
                        dl[dslot][4 + (f - 4)][tid] = dxx + dyy;
                    }
                    else {
                        dl[dslot][7][tid] = dx;
                        dl[dslot][8][tid] = dy;
                        dl[dslot][9][tid] = dxx + dyy;
                    }
                    // in-plane parts of grad(div v): x gets d_xx v_x, y gets d_yy v_y
                    if (f == 1)
                        ch[0][u + 3] += dxx;
                    if (f == 2)
                        ch[1][u + 3] += dyy;
                    if (f == 4)
                        ch[3][u + 3] += dxx;
                    if (f == 5)
                        ch[4][u + 3] += dyy;
                }
                // xy-mixed derivatives of u_x, u_y, A_x, A_y
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const int fx = 1 + 3 * v, fy = 2 + 3 * v;
                    T mxx = T(0), myy = T(0);
#pragma unroll
                    for (int i = 1; i <= 3; ++i) {
                        if constexpr (CANCEL_FREE) {
                            mxx += C.mx[0][i] * ((LD(fx, i, i) - LD(fx, i, -i)) -
                                                 (LD(fx, -i, i) - LD(fx, -i, -i)));
                            myy += C.mx[0][i] * ((LD(fy, i, i) - LD(fy, i, -i)) -
                                                 (LD(fy, -i, i) - LD(fy, -i, -i)));
                        }
                        else {
                            mxx += C.mx[0][i] * ((LD(fx, i, i) + LD(fx, -i, -i)) -
                                                 (LD(fx, i, -i) + LD(fx, -i, i)));
                            myy += C.mx[0][i] * ((LD(fy, i, i) + LD(fy, -i, -i)) -
                                                 (LD(fy, i, -i) + LD(fy, -i, i)));
                        }
                    }
                    ch[3 * v + 0][u + 3] += myy;  // d_xy v_y into (grad div v)_x
                    ch[3 * v + 1][u + 3] += mxx;  // d_xy v_x into (grad div v)_y
                }
            }
            dl[dslot][2][tid] = az_dx_p;
            dl[dslot][3][tid] = az_dy_p;
            dslot = dslot == 2 ? 0 : dslot + 1;
            // pin the carried state: no computation moves across the plane step
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int k = u + 1; k <= u + 6; ++k)
                    asm volatile("" : "+v"(q[f][k]));
#pragma unroll
            for (int c = 0; c < SW_NCH; ++c)
#pragma unroll
                for (int k = u + 1; k <= u + 6; ++k)
                    asm volatile("" : "+v"(ch[c][k]));
#pragma unroll
            for (int k = u + 1; k <= u + 3; ++k) {
                SweepPlane<T>& e = dly[k];
                asm volatile("" : "+v"(e.lr_dx), "+v"(e.lr_dy), "+v"(e.lr_lap), "+v"(e.lr_up));
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    asm volatile("" : "+v"(e.u_dx[i]), "+v"(e.u_dy[i]), "+v"(e.u_lap[i]),
                                 "+v"(e.u_up[i]));
            }

            // plane p + 1 has landed in the other buffer and every wave is done with this one
            plane_barrier();
        };
        plane_step(std::integral_constant<int, 0>{});
        if (it0 + 1 >= niter)
            break;
        plane_step(std::integral_constant<int, 1>{});
        if (it0 + UNR >= niter)
            break;
        // ---- advance the plane-relative state by UNR planes
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                q[f][k] = q[f][k + UNR];
#pragma unroll
        for (int c = 0; c < SW_NCH; ++c)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                ch[c][k] = ch[c][k + UNR];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            dly[k] = dly[k + UNR];
    }
    if (step == 2) {
        if (a.umax2_next != nullptr)
            wave_max_to_global(u2max, a.umax2_next);
        if (a.time_acc != nullptr && blockIdx.x == 0 && tid == 0)
            *a.time_acc += (double)dt;
    }
}

template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY, bool WRAP, int STEP, bool ISO>
__global__ void __launch_bounds__(TX* TY, 2)
    solve_sweep_kernel(SolveArgs<T> a, int zchunk, int ntx, int nty, int nchunks)
{
    sweep_body<T, UPWD, ENTROPY, KB, TX, TY, WRAP, STEP, ISO>(a, zchunk, ntx, nty, nchunks);
}

}  // namespace acamd

namespace acamd {

namespace {

template <typename T, bool UPWD, bool ENTROPY, int KB, bool WRAP>
void sweep_launch(const SolveArgs<T>& a, int zchunk, hipStream_t s)
{
    constexpr int TX = 64, TY = 4;
    const int ntx = a.nx / TX, nty = a.ny / TY, nch = a.nz / zchunk;
    const dim3 tpb(TX, TY, 1);
    const dim3 bpg((unsigned)(ntx * nty * nch), 1, 1);
    // one instantiation per substep: the substep's branches (w read at s > 0, w written at
    // s < 2, max |u| at s = 2) fold away
    if (a.step == 0)
        hipLaunchKernelGGL((solve_sweep_kernel<T, UPWD, ENTROPY, KB, TX, TY, WRAP, 0, true>), bpg,
                           tpb, 0, s, a, zchunk, ntx, nty, nch);
    else if (a.step == 1)
        hipLaunchKernelGGL((solve_sweep_kernel<T, UPWD, ENTROPY, KB, TX, TY, WRAP, 1, true>), bpg,
                           tpb, 0, s, a, zchunk, ntx, nty, nch);
    else
        hipLaunchKernelGGL((solve_sweep_kernel<T, UPWD, ENTROPY, KB, TX, TY, WRAP, 2, true>), bpg,
                           tpb, 0, s, a, zchunk, ntx, nty, nch);
}

template <typename T, bool UPWD, bool ENTROPY>
void sweep_launch_kb(const SolveArgs<T>& a, int kb, int zchunk, hipStream_t s)
{
    auto go = [&](auto KBC) {
        constexpr int KB = decltype(KBC)::value;
        if (a.wrap)
            sweep_launch<T, UPWD, ENTROPY, KB, true>(a, zchunk, s);
        else
            sweep_launch<T, UPWD, ENTROPY, KB, false>(a, zchunk, s);
    };
    if (kb == 0)
        go(std::integral_constant<int, 0>{});
    else if (kb == 8)
        go(std::integral_constant<int, 8>{});
    else if (kb == 16)
        go(std::integral_constant<int, 16>{});
    else
        go(std::integral_constant<int, 32>{});
}

// The three axes share one set of coefficients (equal grid spacing).
template <typename T>
bool isotropic(const Coeffs<T>& C)
{
    for (int o = 0; o < 4; ++o)
        for (int ax = 1; ax < 3; ++ax)
            if (C.d1[ax][o] != C.d1[0][o] || C.d2[ax][o] != C.d2[0][o] ||
                C.up[ax][o] != C.up[0][o] || C.mx[ax][o] != C.mx[0][o])
                return false;
    return true;
}

}  // namespace

bool sweep_shape_fits(TiledShape s, int nx, int ny, int nz)
{
    return sizeof(AcRealT) == 4 && s.kind == 1 && s.tx == 64 && s.ty == 4 && s.zchunk > 0 &&
           nx % 64 == 0 && ny % 4 == 0 && nz % s.zchunk == 0 && nx >= R && ny >= R && nz >= R;
}

template <typename T>
bool sweep_applies(const SolveArgs<T>& a, int kb, TiledShape sh)
{
    if (sizeof(T) != 4 || !sweep_shape_fits(sh, a.nx, a.ny, a.nz) || !isotropic(a.C))
        return false;
    // the kernel addresses field f of every role as slab + f * fstride, and the compensation
    // records as one record of NF entries per cell
    for (int q = 0; q < NF; ++q) {
        const long long o = q * a.fstride;
        if (a.in[q] != a.in[0] + o || a.out[q] != a.out[0] + o || a.w[q] != a.w[0] + o)
            return false;
        if (kb > 0 && (const char*)a.comp[q] != (const char*)a.comp[0] + q * (kb / 8))
            return false;
    }
    return true;
}

template <typename T>
bool launch_solve_sweep(const SolveArgs<T>& a, bool upwd, bool entropy, int kb, TiledShape sh,
                        hipStream_t s)
{
    if constexpr (sizeof(T) != 4) {
        return false;
    }
    else {
        if (!sweep_applies(a, kb, sh))
            return false;
        if (upwd && entropy)
            sweep_launch_kb<T, true, true>(a, kb, sh.zchunk, s);
        else if (upwd)
            sweep_launch_kb<T, true, false>(a, kb, sh.zchunk, s);
        else if (entropy)
            sweep_launch_kb<T, false, true>(a, kb, sh.zchunk, s);
        else
            sweep_launch_kb<T, false, false>(a, kb, sh.zchunk, s);
        return true;
    }
}

template bool launch_solve_sweep<AcRealT>(const SolveArgs<AcRealT>&, bool, bool, int, TiledShape,
                                          hipStream_t);
template bool sweep_applies<AcRealT>(const SolveArgs<AcRealT>&, int, TiledShape);

}  // namespace acamd
