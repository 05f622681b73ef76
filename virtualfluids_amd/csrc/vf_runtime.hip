// Library-wide bookkeeping (ABI version, last error message) and the copy kernel that
// yardsticks HBM bandwidth.
#include <cstdio>
#include <cstring>

#include "vf_common.h"

namespace vfa {

static thread_local char g_last_error[512] = "";

void set_last_error(const char* msg)
{
    std::snprintf(g_last_error, sizeof(g_last_error), "%s", msg);
}

int check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
        return VF_ERR_HIP;
    }
    return VF_OK;
}

typedef unsigned int copy_u4 __attribute__((ext_vector_type(4)));

// one 16-byte element per thread, grid-stride over the rest
__global__ __launch_bounds__(256) void stream_copy_kernel(const copy_u4* __restrict__ src, copy_u4* __restrict__ dst,
                                                          uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

} // namespace vfa

extern "C" int vf_abi_version(void) { return VF_ABI_VERSION; }

extern "C" const char* vf_get_last_error(void) { return vfa::g_last_error; }

extern "C" int vf_stream_copy(const void* src, void* dst, uint64_t bytes, void* stream)
{
    if (bytes == 0) return VF_OK;
    if (!src || !dst || bytes % 16 != 0 || ((uintptr_t)src | (uintptr_t)dst) % 16 != 0) {
        vfa::set_last_error("stream_copy: null pointer, or bytes / pointers not 16-byte aligned");
        return VF_ERR_INVALID_ARGUMENT;
    }
    const uint64_t n = bytes / 16;
    const uint64_t blocks = (n + 255) / 256;
    const unsigned grid = (unsigned)(blocks < (1u << 30) ? blocks : (1u << 30));
    hipLaunchKernelGGL(vfa::stream_copy_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                       static_cast<const vfa::copy_u4*>(src), static_cast<vfa::copy_u4*>(dst), n);
    return vfa::check_launch("stream_copy_kernel execution failed");
}
