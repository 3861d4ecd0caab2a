// ic.hip -- the seeded Fourier-mode initial condition evaluated on the device.
//
// Produces bit for bit what acHostMeshInitFourier (host.cc) writes into a host mesh: the mode
// set and the per-axis cos / sin tables come from the same host code (fourier_tables), and every
// cell evaluates the same separable complex products in the same order, in FP64 with no
// contraction into fused multiply-adds, before the one rounding to AcReal. Large grids (1024^3
// and up) then start in well under a second with no host mesh and no host-to-device copy.
#include "kernels.h"

namespace acamd {

__device__ __forceinline__ unsigned long long mix64_dev(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void __launch_bounds__(256) fourier_ic_kernel(FourierArgs<T> A)
{
#pragma clang fp contract(off)
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    const int j = threadIdx.y + blockIdx.y * blockDim.y;
    const int k = blockIdx.z;
    if (i >= A.nx || j >= A.ny || k >= A.nz)
        return;
    const int I = i + R, J = j + R, K = k + R;
    const long long per = (long long)A.mx + A.my + A.mz;
    const long long c = (long long)I + (long long)A.mx * (J + (long long)A.my * K);
    for (int g = 0; g < 2; ++g) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < A.nmodes; ++m) {
            const long long gm = (long long)g * A.nmodes + m;
            const double* h = A.head + gm * 5;
            const double* t = A.tab + gm * per * 2;
            const double crx = t[2 * I], cix = t[2 * I + 1];
            const double cry = t[2 * (A.mx + J)], ciy = t[2 * (A.mx + J) + 1];
            const double crz = t[2 * (A.mx + A.my + K)], ciz = t[2 * (A.mx + A.my + K) + 1];
            const double pr = h[0], pi = h[1];
            // e^{i phi} e^{i kz z} e^{i ky y}, then the real part of the product with e^{i kx x}
            const double zr = pr * crz - pi * ciz;
            const double zi = pr * ciz + pi * crz;
            const double yr = zr * cry - zi * ciy;
            const double yi = zr * ciy + zi * cry;
            const double cv = yr * crx - yi * cix;
            for (int q = 0; q < 3; ++q)
                acc[q] += h[2 + q] * cv;
        }
        T* const* dst = g == 0 ? A.uu : A.aa;
        for (int q = 0; q < 3; ++q)
            dst[q][c] = (T)acc[q];
    }
    double v = 0.0;
    if (A.lnrho_perturb != 0.0) {
        const double u = (double)(mix64_dev(A.lseed + (unsigned long long)c) >> 11) *
                         (1.0 / 9007199254740992.0);
        v = A.lnrho_perturb * (2.0 * u - 1.0);
    }
    A.lnrho[c] = (T)v;
    A.ss[c] = (T)0;
}

template <typename T>
void launch_fourier_ic(const FourierArgs<T>& A, hipStream_t s)
{
    const dim3 tpb(64, 4, 1);
    const dim3 bpg((unsigned)((A.nx + 63) / 64), (unsigned)((A.ny + 3) / 4), (unsigned)A.nz);
    hipLaunchKernelGGL((fourier_ic_kernel<T>), bpg, tpb, 0, s, A);
}

template void launch_fourier_ic<AcRealT>(const FourierArgs<AcRealT>&, hipStream_t);

}  // namespace acamd
