// FP64 direct sum at sampled targets on caller-owned device buffers, with no device
// allocation: the partial sums live in a fixed file-scope device array (`scratch`), one
// per GPU, whose address is looked up once per device.  For callers that cannot call
// hipMalloc (graph capture, serving loops); hfmm_direct (host buffers, allocating) is the
// entry point it mirrors.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>

#include "../../include/hfmm.h"
#include "common.h"

namespace hfmm {
namespace {

// partial sums of the targets, (re, im) per target: 1 MiB per GPU
static __device__ __align__(256) double2 scratch[HFMM_DIRECT_NOALLOC_MAX_TARGETS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double2* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` per device

// the device address of this device's `scratch`, looked up once per device with the
// device current (the host-side name is a stand-in, not a device address)
double2* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(HFMM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double2*)p;
    }
    return g_scratch_ptr[dev];
}

__global__ __launch_bounds__(256) void scratch_clear_kernel(double2* acc, int64_t nt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nt) acc[i] = make_double2(0.0, 0.0);
}

// grid (target tiles of 256, source chunks): each block stages its chunk's sources 256 at
// a time in LDS (from the interleaved x, y, z array) and adds its partial sums to the
// scratch accumulators with FP64 atomics
__global__ __launch_bounds__(256) void direct_aos_kernel(const double* xyz, const double2* q, int64_t n, double k,
                                                         const int64_t* targets, int64_t nt, int64_t chunk,
                                                         double2* acc) {
    __shared__ double sx[256], sy[256], sz[256], sqr[256], sqi[256];
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = t < nt;
    // a target index outside [0, n) reads point 0 and its result is NaN
    const int64_t tr = valid ? targets[t] : 0;
    const bool in_range = tr >= 0 && tr < n;
    const int64_t ti = in_range ? tr : 0;
    const double x = xyz[3 * ti], y = xyz[3 * ti + 1], z = xyz[3 * ti + 2];
    const int64_t s0 = (int64_t)blockIdx.y * chunk, s1 = min(n, s0 + chunk);
    double ar = 0.0, ai = 0.0;
    for (int64_t base = s0; base < s1; base += 256) {
        const int64_t v = base + threadIdx.x;
        if (v < s1) {
            sx[threadIdx.x] = xyz[3 * v];
            sy[threadIdx.x] = xyz[3 * v + 1];
            sz[threadIdx.x] = xyz[3 * v + 2];
            const double2 qq = q[v];
            sqr[threadIdx.x] = qq.x;
            sqi[threadIdx.x] = qq.y;
        }
        __syncthreads();
        const int nb = (int)min((int64_t)256, s1 - base);
        for (int j = 0; j < nb; ++j) {
            const double dx = x - sx[j], dy = y - sy[j], dz = z - sz[j];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 0.0) {
                double s, c;
                sincos(k * r, &s, &c);
                const double wr = c / r, wi = s / r;
                ar += wr * sqr[j] - wi * sqi[j];
                ai += wr * sqi[j] + wi * sqr[j];
            }
        }
        __syncthreads();
    }
    if (valid) {
        if (!in_range) ar = ai = __builtin_nan("");
        atomicAdd(&acc[t].x, ar);
        atomicAdd(&acc[t].y, ai);
    }
}

__global__ __launch_bounds__(256) void scratch_copy_kernel(const double2* acc, int64_t nt, double2* out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nt) out[i] = acc[i];
}

void check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(HFMM_ERR_DEVICE, std::string("kernel launch failed (") + what + "): " + hipGetErrorString(e));
}

}  // namespace
}  // namespace hfmm

extern "C" int hfmm_direct_noalloc(const double* xyz, int64_t n, const double* q, double k, const int64_t* targets,
                                   int64_t n_targets, double* out, void* stream) {
    using namespace hfmm;
    try {
        if (!xyz || !q || !out || (n_targets > 0 && !targets) || n <= 0 || n_targets < 0)
            fail(HFMM_ERR_ARG, "bad argument");
        if (!std::isfinite(k)) fail(HFMM_ERR_ARG, "k must be a finite real number");
        if (n_targets > HFMM_DIRECT_NOALLOC_MAX_TARGETS)
            fail(HFMM_ERR_ARG, "n_targets exceeds HFMM_DIRECT_NOALLOC_MAX_TARGETS (" +
                                   std::to_string(HFMM_DIRECT_NOALLOC_MAX_TARGETS) + ")");
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) fail(HFMM_ERR_DEVICE, "hipGetDevice failed");
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) fail(HFMM_ERR_DEVICE, "hipGetDeviceProperties failed");
        if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
            fail(HFMM_ERR_DEVICE, std::string("libhfmm is built for gfx950, found ") + prop.gcnArchName);
        double2* acc = scratch_ptr(dev);
        if (n_targets == 0) return HFMM_OK;
        hipStream_t s = (hipStream_t)stream;
        const int64_t tiles = (n_targets + 255) / 256;
        int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(4096 / tiles, (n + 4095) / 4096));
        const int64_t chunk = (n + nchunk - 1) / nchunk;
        nchunk = (n + chunk - 1) / chunk;
        // the accumulators keep the previous call's sums: clear them (a kernel, so that a
        // captured graph holds no memset node)
        hipLaunchKernelGGL(scratch_clear_kernel, dim3((unsigned)tiles), dim3(256), 0, s, acc, n_targets);
        check_launch("scratch clear");
        hipLaunchKernelGGL(direct_aos_kernel, dim3((unsigned)tiles, (unsigned)nchunk), dim3(256), 0, s, xyz,
                           (const double2*)q, n, k, targets, n_targets, chunk, acc);
        check_launch("direct (no allocation)");
        hipLaunchKernelGGL(scratch_copy_kernel, dim3((unsigned)tiles), dim3(256), 0, s, acc, n_targets,
                           (double2*)out);
        check_launch("scratch copy");
        return HFMM_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return HFMM_ERR_MEMORY;
    } catch (...) {
        set_last_error("unknown error");
        return HFMM_ERR_INTERNAL;
    }
}
