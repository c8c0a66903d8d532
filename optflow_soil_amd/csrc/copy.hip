// Every copy the library makes (device to device, host to device, device to host, 2-D crops) is a
// kernel: the runtime's asynchronous copies are never used. A wait on a stream whose only pending
// work was one of those copies has been seen never to return; a kernel copy completes like any
// other kernel. Host data moves through host-mapped pinned memory that the kernels read or write
// directly. All copies are HBM- (or PCIe-) bound streams: 16-byte accesses where both ends are
// 16-byte aligned, bytes otherwise.
#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace ofx {

namespace {

__global__ void k_copy16(uint4 *__restrict__ d, const uint4 *__restrict__ s, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

__global__ void k_copy1(char *__restrict__ d, const char *__restrict__ s, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// rows of `words` 4-byte words; pitches in words
__global__ void k_copy2d(uint32_t *__restrict__ d, size_t dpitch, const uint32_t *__restrict__ s, size_t spitch,
                         size_t words, size_t rows) {
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < words && y < rows) d[y * dpitch + x] = s[y * spitch + x];
}

__global__ void k_twiddles(float2 *__restrict__ tw, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double a = -2.0 * M_PI * (double)e / (double)n;
    tw[e] = make_float2((float)cos(a), (float)sin(a));
}

unsigned blocks_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 4096); }

}  // namespace

void dcopy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0;
    const size_t n16 = aligned ? bytes / 16 : 0;
    if (n16) hipLaunchKernelGGL(k_copy16, dim3(blocks_for(n16)), dim3(256), 0, s, (uint4 *)dst, (const uint4 *)src, n16);
    const size_t done = n16 * 16;
    if (done < bytes)
        hipLaunchKernelGGL(k_copy1, dim3(blocks_for(bytes - done)), dim3(256), 0, s, (char *)dst + done,
                           (const char *)src + done, bytes - done);
    LAUNCHCHECK();
}

void dcopy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height, hipStream_t s) {
    if (!width || !height) return;
    if ((width | dpitch | spitch) & 3) throw Error(OFX_EINVAL, "dcopy2d: rows and pitches must be whole 4-byte words");
    const size_t words = width / 4;
    hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((words + 255) / 256), (unsigned)height), dim3(256), 0, s,
                       (uint32_t *)dst, dpitch / 4, (const uint32_t *)src, spitch / 4, words, height);
    LAUNCHCHECK();
}

void make_twiddles(float2 *tw, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, s, tw, n);
    LAUNCHCHECK();
}

}  // namespace ofx
