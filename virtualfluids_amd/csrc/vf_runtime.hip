// Library-wide bookkeeping (ABI version, last error message) and the copy and in-place kernels
// that yardstick HBM bandwidth.
#include <atomic>
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

size_t lds_reserve_for(int k)
{
    if (k <= 0) return 0;
    static std::atomic<int> cached[64]; // LDS bytes per CU of each device, 0 until queried
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (dev >= 0 && dev < 64) lds = cached[dev].load(std::memory_order_relaxed);
    if (lds <= 0) {
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || lds <= 0)
            return 0;
        if (dev >= 0 && dev < 64) cached[dev].store(lds, std::memory_order_relaxed);
    }
    // whole KiB below the share of one workgroup, less 1 KiB for the kernel's own LDS words and
    // the allocation granularity
    const long bytes = (long)(lds / k / 1024) * 1024 - 1024;
    return bytes > 0 ? (size_t)bytes : 0;
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

// In-place read-modify-write of 27 streams: every lane reads one element of each stream and
// writes it back into the opposite stream at the same offset (d <-> inv(d), stream 0 in place),
// the memory pattern of the in-place collision without its arithmetic. A workgroup covers 256
// consecutive elements of every stream, as the FP32 row kernel covers one 256-node row. Dynamic
// LDS reserved at launch (and never touched) caps the resident workgroups per CU.
__global__ __launch_bounds__(256) void stream_inplace27_kernel(float* f, uint64_t n)
{
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    float v[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) v[d] = __builtin_nontemporal_load(f + d * n + k);
#pragma unroll
    for (int d = 0; d < 27; ++d) __builtin_nontemporal_store(v[inv(d)], f + d * n + k);
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

extern "C" int vf_stream_inplace27(void* buf, uint64_t streamBytes, int32_t workgroupsPerCU, void* stream)
{
    if (streamBytes == 0) return VF_OK;
    if (!buf || streamBytes % 1024 != 0 || (uintptr_t)buf % 16 != 0) {
        vfa::set_last_error("stream_inplace27: null buffer, or streamBytes not a multiple of 1024");
        return VF_ERR_INVALID_ARGUMENT;
    }
    const uint64_t n = streamBytes / 4;
    const uint64_t blocks = n / 256;
    if (blocks >= (1ull << 24)) { // a launch holds fewer than 2^32 work-items
        vfa::set_last_error("stream_inplace27: streams longer than 2^32 - 256 elements");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (workgroupsPerCU < 0 || workgroupsPerCU > 32) {
        vfa::set_last_error("stream_inplace27: workgroupsPerCU must lie in [0, 32]");
        return VF_ERR_INVALID_ARGUMENT;
    }
    const size_t lds = vfa::lds_reserve_for(workgroupsPerCU);
    hipLaunchKernelGGL(vfa::stream_inplace27_kernel, dim3((uint32_t)blocks), dim3(256), lds, (hipStream_t)stream,
                       static_cast<float*>(buf), n);
    return vfa::check_launch("stream_inplace27_kernel execution failed");
}
