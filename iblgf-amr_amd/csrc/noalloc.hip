// Fixed device scratch of the allocation-free solve entry point
// (iblgf_gpu_apply_lgf_noalloc).
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy, at its own address, from module load until the process ends.
// In host code its name is only a stand-in, so the device address is looked
// up once per device (hipGetSymbolAddress with that device current, at context
// creation) and only that address is ever handed to kernels.
#include "noalloc.hpp"

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

namespace iblgf_amd {
namespace {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex           g_lock;
std::map<int, void*> g_addr;  // device -> address of its scratch

__global__ __launch_bounds__(256) void k_zero16(uint4* __restrict__ p, size_t n)
{
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = z;
}

}  // namespace

void* noalloc_scratch_init(int device)
{
    std::lock_guard<std::mutex> g(g_lock);
    const auto                  it = g_addr.find(device);
    if (it != g_addr.end()) return it->second;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device)
        throw std::runtime_error("noalloc scratch: device " + std::to_string(device) + " is not current");
    void*            ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw std::runtime_error(std::string("noalloc scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_addr[device] = ptr;
    return ptr;
}

void* noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> g(g_lock);
    const auto                  it = g_addr.find(device);
    return it == g_addr.end() ? nullptr : it->second;
}

void noalloc_clear(hipStream_t s, void* p, size_t bytes)
{
    if (bytes > kNoallocScratchBytes || bytes % 16 != 0)
        throw std::invalid_argument("noalloc_clear: size outside the scratch or not a multiple of 16 bytes");
    const size_t n = bytes / 16;
    if (n == 0) return;
    const size_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_zero16, dim3(static_cast<unsigned>(blocks < 4096 ? blocks : 4096)), dim3(256), 0, s,
        static_cast<uint4*>(p), n);
}

}  // namespace iblgf_amd
