// Device memory, streams and events for callers that have no GPU runtime of their own
// (the Python binding, the batched runner and the serving loop). Everything goes
// through the one HIP runtime libminifem is linked against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "minifem.h"

namespace {

// Streaming copy, four 16-byte loads in flight per lane, then four stores: the HBM
// bandwidth reference the roofline and the byte counters are calibrated against.
__global__ void __launch_bounds__(256)
    k_hbm_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, int64_t n16)
{
    constexpr int U = 4;
    const int64_t stride = (int64_t)gridDim.x * 256;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + (U - 1) * stride < n16; i += U * stride) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; u++) dst[i + u * stride] = v[u];
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

// Random gather of 8-byte values: dst[i] = src[idx[i]]. Measurement only: the byte
// counters' scale for scattered reads, which move whole DRAM sectors per value.
__global__ void __launch_bounds__(256)
    k_gather(double *__restrict__ dst, const double *__restrict__ src,
             const int64_t *__restrict__ idx, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}

}  // namespace

extern "C" int mf_gather_probe(double *dst, const double *src, const int64_t *idx, int64_t n,
                               void *stream)
{
    if (n < 0 || (n > 0 && (!dst || !src || !idx))) return MF_ERR_ARG;
    if (n == 0) return MF_OK;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dst, src, idx, n);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_hbm_copy(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return MF_OK;
    if (!dst || !src || (bytes & 15) || (reinterpret_cast<uintptr_t>(dst) & 15) ||
        (reinterpret_cast<uintptr_t>(src) & 15))
        return MF_ERR_ARG;
    const int64_t n16 = (int64_t)(bytes / 16);
    const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(k_hbm_copy, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), reinterpret_cast<uint4 *>(dst),
                       reinterpret_cast<const uint4 *>(src), n16);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_dev_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return MF_ERR_ARG;
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    return hipMalloc(ptr, bytes) == hipSuccess ? MF_OK : MF_ERR_NOMEM;
}

extern "C" int mf_dev_free(void *ptr)
{
    if (!ptr) return MF_OK;
    return hipFree(ptr) == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return MF_OK;
    if (!dst || !src) return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return MF_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return MF_OK;
    if (!dst || !src) return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) return MF_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return MF_OK;
    if (!dst || !src) return MF_ERR_ARG;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                          reinterpret_cast<hipStream_t>(stream)) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_stream_create(void **stream)
{
    if (!stream) return MF_ERR_ARG;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MF_ERR_HIP;
    *stream = s;
    return MF_OK;
}

extern "C" int mf_stream_destroy(void *stream)
{
    if (!stream) return MF_OK;
    return hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)) == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_stream_sync(void *stream)
{
    return hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)) == hipSuccess ? MF_OK
                                                                                      : MF_ERR_HIP;
}

extern "C" int mf_device_sync(void)
{
    return hipDeviceSynchronize() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_event_create(void **event)
{
    if (!event) return MF_ERR_ARG;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return MF_ERR_HIP;
    *event = e;
    return MF_OK;
}

extern "C" int mf_event_destroy(void *event)
{
    if (!event) return MF_OK;
    return hipEventDestroy(reinterpret_cast<hipEvent_t>(event)) == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_event_record(void *event, void *stream)
{
    if (!event) return MF_ERR_ARG;
    return hipEventRecord(reinterpret_cast<hipEvent_t>(event),
                          reinterpret_cast<hipStream_t>(stream)) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!start || !stop || !ms) return MF_ERR_ARG;
    if (hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)) != hipSuccess) return MF_ERR_HIP;
    return hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                               reinterpret_cast<hipEvent_t>(stop)) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_host_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return MF_ERR_ARG;
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    return hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess ? MF_OK : MF_ERR_NOMEM;
}

extern "C" int mf_host_free(void *ptr)
{
    if (!ptr) return MF_OK;
    return hipHostFree(ptr) == hipSuccess ? MF_OK : MF_ERR_HIP;
}

extern "C" int mf_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return MF_OK;
    if (!dst || !src) return MF_ERR_ARG;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                          reinterpret_cast<hipStream_t>(stream)) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return MF_OK;
    if (!dst || !src) return MF_ERR_ARG;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                          reinterpret_cast<hipStream_t>(stream)) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_event_query(void *event, int *done)
{
    if (!event || !done) return MF_ERR_ARG;
    const hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(event));
    if (e == hipSuccess) {
        *done = 1;
        return MF_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return MF_OK;
    }
    return MF_ERR_HIP;
}

extern "C" int mf_stream_wait_event(void *stream, void *event)
{
    if (!event) return MF_ERR_ARG;
    return hipStreamWaitEvent(reinterpret_cast<hipStream_t>(stream),
                              reinterpret_cast<hipEvent_t>(event), 0) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_graph_begin(void *stream)
{
    if (!stream) return MF_ERR_ARG;  // capture needs a created stream
    return hipStreamBeginCapture(reinterpret_cast<hipStream_t>(stream),
                                 hipStreamCaptureModeRelaxed) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_graph_end(void *stream, void **graph)
{
    if (!stream || !graph) return MF_ERR_ARG;
    *graph = nullptr;
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(reinterpret_cast<hipStream_t>(stream), &g) != hipSuccess || !g)
        return MF_ERR_HIP;
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return MF_ERR_HIP;
    *graph = x;
    return MF_OK;
}

extern "C" int mf_graph_launch(void *graph, void *stream)
{
    if (!graph) return MF_ERR_ARG;
    return hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph),
                          reinterpret_cast<hipStream_t>(stream)) == hipSuccess
               ? MF_OK
               : MF_ERR_HIP;
}

extern "C" int mf_graph_destroy(void *graph)
{
    if (!graph) return MF_OK;
    return hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)) == hipSuccess ? MF_OK
                                                                                       : MF_ERR_HIP;
}
