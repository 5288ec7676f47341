// Launch-floor probes for the measurement harness (bench.py --form floor):
// kernels with the executor's grid and block shape that do (almost) nothing,
// so their rocprofv3 duration is the floor every executor launch pays.
//   kind 0: empty body
//   kind 1: every wave issues `depth` (1..4) dependent scalar loads from `buf`
//           (the first from a per-wave cache line, each next address taken
//           from the loaded word), the latency chain a cold executor wave
//           walks before its first B load (descriptor -> record -> values);
//           `buf` holds zeros, so every address stays inside the buffer.
#include <hip/hip_runtime.h>

#include "srt.h"
#include "srt_internal.h"

namespace srt {
namespace {

__global__ void floor_empty_kernel() {}

template <int DEPTH>
__global__ void floor_chain_kernel(const int32_t *__restrict__ buf, int64_t words, int32_t *__restrict__ sink) {
    constexpr int depth = DEPTH;
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * 32;  // one 128-byte line per wave and step
    int64_t at = ((int64_t)wave * 32) % words;
    int32_t v = 0;
    for (int d = 0; d < depth; ++d) {
        v = __builtin_amdgcn_readfirstlane(buf[at]);
        at = (at + stride + (int64_t)(uint32_t)v) % words;
    }
    if (v == 0x7fffffff && threadIdx.x == 0) sink[0] = v;  // keeps the loads; never true for a zero buffer
}

}  // namespace
}  // namespace srt

extern "C" int srt_probe_launch_floor(int kind, int grid, int block, const void *buf, int64_t buf_bytes,
                                      int depth, void *sink, srt_stream_t stream) {
    using namespace srt;
    if (grid < 1 || block < 64 || block > 1024 || block % 64)
        return fail(SRT_ERR_INVALID, "probe: grid >= 1, block a multiple of 64 in [64, 1024]");
    if (kind == 0) {
        hipLaunchKernelGGL(floor_empty_kernel, dim3(grid), dim3(block), 0, (hipStream_t)stream);
    } else if (kind == 1) {
        if (!buf || !sink || buf_bytes < 128 || depth < 1 || depth > 4)
            return fail(SRT_ERR_INVALID, "probe: chain needs a buffer of >= 128 bytes, a sink and depth in [1, 4]");
        const int32_t *b = (const int32_t *)buf;
        const int64_t w = buf_bytes / 4;
        int32_t *k = (int32_t *)sink;
        hipStream_t st = (hipStream_t)stream;
        // one instantiation per depth: rocprofv3 reports each under its own name
        if (depth == 1) hipLaunchKernelGGL(floor_chain_kernel<1>, dim3(grid), dim3(block), 0, st, b, w, k);
        if (depth == 2) hipLaunchKernelGGL(floor_chain_kernel<2>, dim3(grid), dim3(block), 0, st, b, w, k);
        if (depth == 3) hipLaunchKernelGGL(floor_chain_kernel<3>, dim3(grid), dim3(block), 0, st, b, w, k);
        if (depth == 4) hipLaunchKernelGGL(floor_chain_kernel<4>, dim3(grid), dim3(block), 0, st, b, w, k);
    } else {
        return fail(SRT_ERR_INVALID, "probe: kind must be 0 or 1");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SRT_OK : fail(SRT_ERR_HIP, hipGetErrorString(e));
}
