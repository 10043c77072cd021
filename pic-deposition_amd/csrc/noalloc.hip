// Allocation-free deposit (pic_hip_deposit_noalloc, include/pic_hip.h): the
// global-atomic deposit of pic_hip_deposit (backend GPU_ATOMICS) on
// caller-owned device arrays, for callers that cannot allocate device memory
// (work under HIP graph capture, serving loops).  Its only scratch memory, the
// FP64 accumulation grid and a status word, is the file-scope device array
// `scratch` below, one copy per device, resolved to a device address once per
// device with hipGetSymbolAddress.

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/pic_hip.h"
#include "api_util.hpp"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

using namespace pichip;

namespace {

// FP64 accumulation grid of up to PIC_NOALLOC_MAX_CELLS cells, followed by
// one 256-byte line holding the status word of the Esirkepov long-move path.
constexpr size_t kStatusWords = 32;  // doubles: one 256-byte line
static __device__ __align__(256) double scratch[PIC_NOALLOC_MAX_CELLS + kStatusWords];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device (null: not resolved yet)

// Zeroes the first n cells of the accumulation grid and the status word (a
// kernel, not a memset, so that the call can be captured into a graph).
__global__ __launch_bounds__(256) void noalloc_clear_kernel(double* __restrict__ acc, size_t n,
                                                            int* __restrict__ status) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = 0.0;
    if (i == 0) *status = 0;
}

// A particle whose Esirkepov move was longer than the walk supports left the
// status word set: every cell of rho becomes NaN, so the failure shows in the
// caller's output (an asynchronous call has no other way to report it).
__global__ __launch_bounds__(256) void noalloc_poison_kernel(float* __restrict__ rho, size_t n,
                                                             const int* __restrict__ status) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && *status != 0) rho[i] = __builtin_nanf("");
}

// Device address of this device's `scratch`, resolved on the first call on
// the device (with that device current).
pic_status scratch_for_current_device(double** out) {
    int dev = 0;
    PIC_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PIC_ERR_UNSUPPORTED, "device index beyond the supported 64");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PIC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return PIC_OK;
}

}  // namespace

extern "C" {

pic_status pic_hip_deposit_noalloc(const pic_domain* domain, const float* x, const float* y, const float* vx,
                                   const float* vy, const float* q, size_t n, int scheme, double dt, float* rho,
                                   void* stream) {
    if (!valid_domain(domain))
        return fail(PIC_ERR_INVALID_ARGUMENT, "invalid domain (need Nx, Ny > 0, L > 0, Nx*Ny < 2^30)");
    const size_t G = (size_t)domain->Nx * (size_t)domain->Ny;
    if (G > (size_t)PIC_NOALLOC_MAX_CELLS)
        return fail(PIC_ERR_INVALID_ARGUMENT, "grid of " + std::to_string(G) + " cells exceeds the " +
                                                  std::to_string(PIC_NOALLOC_MAX_CELLS) +
                                                  "-cell accumulation grid of pic_hip_deposit_noalloc");
    if (scheme < 0 || scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "unknown scheme (0 NGP, 1 CIC, 2 TSC, 3 Esirkepov)");
    if (!rho) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho");
    if (n > 0 && (!x || !y || !q)) return fail(PIC_ERR_INVALID_ARGUMENT, "x, y and q are required");
    if (n > 0 && scheme == PIC_SCHEME_ESIRKEPOV && (!vx || !vy))
        return fail(PIC_ERR_INVALID_ARGUMENT, "the Esirkepov scheme needs vx and vy");
    if (n > ((size_t)1 << 40)) return fail(PIC_ERR_INVALID_ARGUMENT, "particle count too large");
    if (!std::isfinite(dt)) return fail(PIC_ERR_INVALID_ARGUMENT, "dt must be finite");
    double* acc = nullptr;
    const pic_status st = scratch_for_current_device(&acc);
    if (st != PIC_OK) return st;
    int* status = reinterpret_cast<int*>(acc + PIC_NOALLOC_MAX_CELLS);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(noalloc_clear_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, acc, G, status);
    DepositArgs a{};
    a.x = x;
    a.y = y;
    a.vx = vx;
    a.vy = vy;
    a.q = q;
    a.n = (int64_t)n;
    a.cell_off = nullptr;
    a.dom = make_dev_domain(*domain);
    a.scheme = scheme;
    a.backend = PIC_BACKEND_GPU_ATOMICS;
    a.dt = (float)dt;
    a.rho = rho;
    a.partial = nullptr;
    a.acc = acc;
    a.status = status;
    a.stream = s;
    PIC_HIP_TRY(launch_deposit(a));
    hipLaunchKernelGGL(noalloc_poison_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, rho, G, status);
    PIC_HIP_TRY(hipGetLastError());
    return PIC_OK;
}

}  // extern "C"
