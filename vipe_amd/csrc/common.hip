// ABI-level helpers of libvipe_hip.
#include "vh_common.hpp"

namespace {
__global__ void fill_i32_kernel(int32_t* out, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = v;
}
}  // namespace

extern "C" {

const char* vh_last_error(void) { return vh::last_error_buf(); }

int vh_abi_version(void) { return 1; }

int vh_fill_i32(int32_t* out, int64_t n, int32_t value, void* stream) {
    if (n < 0 || (n > 0 && out == nullptr)) return vh::fail(VH_ERR_ARG, "fill_i32: bad arguments");
    if (n == 0) return VH_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out, n, value);
    VH_LAUNCH_CHECK("fill_i32");
    return VH_OK;
}

}  // extern "C"
