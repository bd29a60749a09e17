// sdp_runtime.hip -- device runtime of the C-ABI: initialisation, error
// reporting, memory, streams and events.
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "sdp_internal.h"

namespace {
thread_local char g_error[512] = "";
std::once_flag g_init_once;
int g_init_status = SDP_OK;
hipStream_t g_default_stream = nullptr;
int g_degrid_mode = SDP_DEGRID_NEAREST;
int g_dft_mode = SDP_DFT_FP64_PHASE;

// Zero fill with a kernel (not hipMemsetAsync) so that callers can capture
// it in a graph.
__global__ void zero_kernel(uint4 *p, size_t n16, unsigned char *tail, size_t ntail)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}
// Plain 16-byte copy (graph safe, and the reference copy for HBM counters
// and bandwidth): four loads in flight per thread before their stores.
__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}
}  // namespace

namespace sdp {
int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    return code;
}

hipStream_t stream_of(sdp_stream s)
{
    return s ? (hipStream_t)s : g_default_stream;
}
}  // namespace sdp

extern "C" {

int sdp_init(void)
{
    std::call_once(g_init_once, [] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n < 1) {
            g_init_status = sdp::set_error(SDP_E_HIP, "no HIP device available (%s)",
                                           hipGetErrorString(e));
            return;
        }
        e = hipStreamCreateWithFlags(&g_default_stream, hipStreamNonBlocking);
        if (e != hipSuccess)
            g_init_status = sdp::set_error(SDP_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    });
    return g_init_status;
}

const char *sdp_last_error(void) { return g_error; }

int sdp_set_last_error(int code, const char *message)
{
    return sdp::set_error(code, "%s", message ? message : "");
}

int sdp_device_malloc(void **ptr, size_t bytes)
{
    SDP_REQUIRE(ptr, "sdp_device_malloc: null ptr");
    *ptr = nullptr;
    if (bytes == 0) return SDP_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess)
        return sdp::set_error(SDP_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return SDP_OK;
}

int sdp_device_free(void *ptr)
{
    if (ptr) SDP_HIP_TRY(hipFree(ptr));
    return SDP_OK;
}

int sdp_memcpy_htod(void *dst, const void *src, size_t bytes, sdp_stream s)
{
    if (!bytes) return SDP_OK;
    SDP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, sdp::stream_of(s)));
    SDP_HIP_TRY(hipStreamSynchronize(sdp::stream_of(s)));
    return SDP_OK;
}

int sdp_memcpy_dtoh(void *dst, const void *src, size_t bytes, sdp_stream s)
{
    if (!bytes) return SDP_OK;
    SDP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, sdp::stream_of(s)));
    SDP_HIP_TRY(hipStreamSynchronize(sdp::stream_of(s)));
    return SDP_OK;
}

int sdp_memcpy_dtod(void *dst, const void *src, size_t bytes, sdp_stream s)
{
    if (!bytes || dst == src) return SDP_OK;
    SDP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, sdp::stream_of(s)));
    return SDP_OK;
}

int sdp_memset_zero(void *dst, size_t bytes, sdp_stream s)
{
    if (!bytes) return SDP_OK;
    SDP_REQUIRE(((uintptr_t)dst & 15) == 0, "sdp_memset_zero: pointer must be 16-byte aligned");
    size_t n16 = bytes / 16;
    unsigned int blocks = n16 ? sdp_div_up(n16, 256) : 1;
    if (blocks > 65536) blocks = 65536;
    zero_kernel<<<blocks, 256, 0, sdp::stream_of(s)>>>((uint4 *)dst, n16, (unsigned char *)dst + n16 * 16,
                                                      bytes - n16 * 16);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_device_copy(void *dst, const void *src, size_t bytes, sdp_stream s)
{
    if (!bytes || dst == src) return SDP_OK;
    SDP_REQUIRE(((uintptr_t)dst & 15) == 0 && ((uintptr_t)src & 15) == 0 && (bytes & 15) == 0,
                "sdp_device_copy: pointers and size must be multiples of 16 bytes");
    const size_t n16 = bytes / 16;
    unsigned int blocks = sdp_div_up(n16, 256);
    if (blocks > 16384) blocks = 16384;
    copy_kernel<<<blocks, 256, 0, sdp::stream_of(s)>>>((const uint4 *)src, (uint4 *)dst, n16);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_stream_create(sdp_stream *s)
{
    hipStream_t h;
    SDP_HIP_TRY(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = (sdp_stream)h;
    return SDP_OK;
}

int sdp_stream_destroy(sdp_stream s)
{
    if (s) SDP_HIP_TRY(hipStreamDestroy((hipStream_t)s));
    return SDP_OK;
}

int sdp_stream_sync(sdp_stream s)
{
    SDP_HIP_TRY(hipStreamSynchronize(sdp::stream_of(s)));
    return SDP_OK;
}

int sdp_device_sync(void)
{
    SDP_HIP_TRY(hipDeviceSynchronize());
    return SDP_OK;
}

int sdp_event_create(sdp_event *e)
{
    hipEvent_t h;
    SDP_HIP_TRY(hipEventCreate(&h));
    *e = (sdp_event)h;
    return SDP_OK;
}

int sdp_event_destroy(sdp_event e)
{
    if (e) SDP_HIP_TRY(hipEventDestroy((hipEvent_t)e));
    return SDP_OK;
}

int sdp_event_record(sdp_event e, sdp_stream s)
{
    SDP_HIP_TRY(hipEventRecord((hipEvent_t)e, sdp::stream_of(s)));
    return SDP_OK;
}

int sdp_event_sync(sdp_event e)
{
    SDP_HIP_TRY(hipEventSynchronize((hipEvent_t)e));
    return SDP_OK;
}

int sdp_event_elapsed_ms(sdp_event a, sdp_event b, float *ms)
{
    SDP_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return SDP_OK;
}

int sdp_graph_begin_capture(sdp_stream s)
{
    SDP_HIP_TRY(hipStreamBeginCapture(sdp::stream_of(s), hipStreamCaptureModeThreadLocal));
    return SDP_OK;
}

int sdp_graph_end_capture(sdp_stream s, sdp_graph *graph)
{
    SDP_REQUIRE(graph, "sdp_graph_end_capture: null output");
    hipGraph_t g = nullptr;
    SDP_HIP_TRY(hipStreamEndCapture(sdp::stream_of(s), &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return sdp::set_error(SDP_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    *graph = (sdp_graph)ex;
    return SDP_OK;
}

int sdp_graph_launch(sdp_graph graph, sdp_stream s)
{
    SDP_HIP_TRY(hipGraphLaunch((hipGraphExec_t)graph, sdp::stream_of(s)));
    return SDP_OK;
}

int sdp_graph_destroy(sdp_graph graph)
{
    if (graph) SDP_HIP_TRY(hipGraphExecDestroy((hipGraphExec_t)graph));
    return SDP_OK;
}

void sdp_set_degrid_mode(int mode) { g_degrid_mode = mode == SDP_DEGRID_WEIGHTED ? mode : SDP_DEGRID_NEAREST; }
int sdp_get_degrid_mode(void) { return g_degrid_mode; }
void sdp_set_dft_mode(int mode)
{
    if (mode == SDP_DFT_REFERENCE_FP32 || mode == SDP_DFT_FP64_PHASE) g_dft_mode = mode;
}
int sdp_get_dft_mode(void) { return g_dft_mode; }

}  // extern "C"
