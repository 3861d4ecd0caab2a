// noalloc.hip -- the reductions of acDeviceReduceScal / Vec / VecScal over caller-owned device
// buffers with no allocation: for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops). The result stays in device memory; nothing synchronises the host.
//
// The partial results of the two-stage reduction live in `scratch`, a fixed array in the code
// object's device memory (one copy per GPU, held until the process ends). Host code never passes
// the array's name to a kernel: on the host it is only a stand-in symbol. Its device address is
// looked up once per GPU with hipGetSymbolAddress, with that GPU current, and that address is
// what the kernels receive. Calls that overlap on one GPU share the array and race; the partials
// are all written by the first stage before the second reads them, so nothing needs clearing
// between calls.
#include <cmath>
#include <cstdio>
#include <mutex>

#include "astaroth.h"
#include "kernels.h"

namespace {

using acamd::AcRealT;

// Partials of the first stage (at most acamd::reduce_num_partials(), which caps the grid of that
// stage at 1024 workgroups) plus the raw FP64 result of the second.
constexpr int kNoAllocPartials = 1024;
static __device__ __align__(256) double scratch[kNoAllocPartials + 1];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch_ptr[kMaxDevices] = {};

AcResult noalloc_fail(const char* msg)
{
    fprintf(stderr, "astaroth_amd: acReduceNoAlloc: %s\n", msg);
    return AC_FAILURE;
}

// Device address of this GPU's `scratch` (looked up on the first call with the GPU current).
double* scratch_ptr(int dev)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
            return nullptr;
        g_scratch_ptr[dev] = (double*)p;
    }
    return g_scratch_ptr[dev];
}

// The post-processing of the host-side entry points (sqrt(v / n) for AC_RMS, ref
// src/core/device.cc:1403-1443), applied on the device to the raw result, then rounded to AcReal.
__global__ void reduce_finish_kernel(const double* raw, int rms, double ncells, AcRealT* result)
{
    double v = *raw;
    if (rms)
        v = sqrt(v / ncells);
    *result = (AcRealT)v;
}

}  // namespace

extern "C" AcResult acReduceNoAlloc(const AcReduction reduction, const AcReal* const* vtxbufs,
                                    const int num_vtxbufs, const int nx, const int ny,
                                    const int nz, const int post_process, void* stream,
                                    AcReal* result)
{
    if (vtxbufs == nullptr || result == nullptr)
        return noalloc_fail("null buffer list or result");
    if (nx < 1 || ny < 1 || nz < 1)
        return noalloc_fail("grid extents must be positive");
    int map;
    switch (num_vtxbufs) {
    case 1: map = reduction.map_vtxbuf_single; break;
    case 3: map = reduction.map_vtxbuf_vec; break;
    case 4: map = reduction.map_vtxbuf_vec_scal; break;
    default: return noalloc_fail("num_vtxbufs must be 1 (scalar), 3 (vector) or 4 (vector + scalar)");
    }
    if (map == AC_NULL_KERNEL || reduction.reduce_op == AC_REDUCE_OP_NO_REDUCE)
        return noalloc_fail("reduction type not defined for this number of buffers");
    const AcReal* bufs[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int q = 0; q < num_vtxbufs; ++q) {
        if (vtxbufs[q] == nullptr)
            return noalloc_fail("null vertex buffer");
        bufs[q] = vtxbufs[q];
    }
    const int parts = acamd::reduce_num_partials(nx, ny, nz);
    if (parts > kNoAllocPartials)
        return noalloc_fail("grid needs more partial results than the fixed scratch holds");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return noalloc_fail("no current GPU");
    double* sp = scratch_ptr(dev);
    if (sp == nullptr)
        return noalloc_fail("hipGetSymbolAddress of the scratch array failed");
    hipStream_t s = (hipStream_t)stream;
    acamd::launch_reduce<AcRealT>(bufs, map, (int)reduction.reduce_op, nx, ny, nz, sp,
                                  kNoAllocPartials, sp + kNoAllocPartials, s);
    const int rms = post_process && reduction.post_processing_op == AC_RMS;
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(1), 0, s, (const double*)(sp + kNoAllocPartials),
                       rms, (double)nx * ny * nz, (AcRealT*)result);
    if (hipGetLastError() != hipSuccess)
        return noalloc_fail("kernel launch failed");
    return AC_SUCCESS;
}
