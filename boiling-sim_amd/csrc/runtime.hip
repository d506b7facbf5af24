// Boundary helpers: the FFI smoke kernel (crates/cuda-kernels/src/scale.cu:11-34),
// error strings and device queries for the Python shim.
#include "boilsim_hip.h"
#include "common.hpp"
#include "precond_api.hpp"

#include <stdio.h>
#include <string.h>

namespace bs {

__global__ __launch_bounds__(256) void scale_inplace_kernel(float* a, float s, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) a[t] = a[t] * s;
}

}  // namespace bs

// In-place a[i] *= scale, synchronous (scale.cu:22-34).
extern "C" int bs_scale_inplace_launch(float* d_a, int n, float scale) {
    if (n <= 0) return hipSuccess;
    bs::scale_inplace_kernel<<<(n + 255) / 256, 256>>>(d_a, scale, n);
    BS_RETURN_IF_ERR(hipGetLastError());
    return (int)hipDeviceSynchronize();
}

extern "C" const char* bs_error_string(int code) { return hipGetErrorString((hipError_t)code); }

// Selects device `device_id` for the calling thread and writes a one-line
// description into buf.  Returns a hipError_t.
extern "C" int bs_device_open(int device_id, char* buf, int buflen) {
    int count = 0;
    BS_RETURN_IF_ERR(hipGetDeviceCount(&count));
    if (device_id < 0 || device_id >= count) return (int)hipErrorInvalidDevice;
    BS_RETURN_IF_ERR(hipSetDevice(device_id));
    // The no-allocation solver's scratch address on this device.
    BS_RETURN_IF_ERR((hipError_t)bs::scratch_setup(nullptr));
    hipDeviceProp_t prop;
    BS_RETURN_IF_ERR(hipGetDeviceProperties(&prop, device_id));
    if (buf && buflen > 0)
        snprintf(buf, (size_t)buflen, "HIP device %d ready: %s (%s), %d CUs, %.1f GiB", device_id,
                 prop.name, prop.gcnArchName, prop.multiProcessorCount,
                 (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    return hipSuccess;
}

extern "C" int bs_device_synchronize(void) { return (int)hipDeviceSynchronize(); }

extern "C" int bs_abi_version(void) { return 2; }
