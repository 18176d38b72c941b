// Runtime half of the C ABI: errors, device query, memory, copies, streams, events.
#include "ph3_common.h"
#include "../../include/pyheat3d_hip.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

namespace ph3host {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int check(hipError_t e, const char* what) {
    set_error("%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    return e == hipErrorOutOfMemory ? PH3_ENOMEM : PH3_EHIP;
}

}  // namespace ph3host

using ph3host::set_error;

extern "C" {

const char* ph3_last_error(void) { return ph3host::g_err; }

int ph3_version(void) { return 10000; }  // 1.0.0

int ph3_device_count(int* n) {
    if (!n) { set_error("ph3_device_count: n is NULL"); return PH3_EINVAL; }
    *n = 0;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) { (void)hipGetLastError(); return PH3_OK; }
    PH3_CHECK(e);
    *n = c;
    return PH3_OK;
}

int ph3_device_info(int dev, char* name, int name_len, char* arch, int arch_len, int* cu_count,
                    size_t* total_mem, int* wave_size, int* lds_per_cu, int* mem_clock_khz,
                    int* bus_width_bits) {
    hipDeviceProp_t p;
    PH3_CHECK(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) { strncpy(name, p.name, name_len - 1); name[name_len - 1] = 0; }
    if (arch && arch_len > 0) { strncpy(arch, p.gcnArchName, arch_len - 1); arch[arch_len - 1] = 0; }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (total_mem) *total_mem = p.totalGlobalMem;
    if (wave_size) *wave_size = p.warpSize;
    if (lds_per_cu) *lds_per_cu = (int)p.maxSharedMemoryPerMultiProcessor;
    if (mem_clock_khz) *mem_clock_khz = p.memoryClockRate;
    if (bus_width_bits) *bus_width_bits = p.memoryBusWidth;
    return PH3_OK;
}

int ph3_set_device(int dev) { PH3_CHECK(hipSetDevice(dev)); return PH3_OK; }

int ph3_device_sync(void) { PH3_CHECK(hipDeviceSynchronize()); return PH3_OK; }

int ph3_mem_info(size_t* free_bytes, size_t* total_bytes) {
    size_t f = 0, t = 0;
    PH3_CHECK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return PH3_OK;
}

int ph3_malloc(void** ptr, size_t bytes) {
    if (!ptr) { set_error("ph3_malloc: ptr is NULL"); return PH3_EINVAL; }
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    PH3_CHECK(hipMalloc(ptr, bytes));
    return PH3_OK;
}

int ph3_free(void* ptr) {
    if (ptr) PH3_CHECK(hipFree(ptr));
    return PH3_OK;
}

int ph3_host_alloc(void** ptr, size_t bytes) {
    if (!ptr) { set_error("ph3_host_alloc: ptr is NULL"); return PH3_EINVAL; }
    *ptr = nullptr;
    PH3_CHECK(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return PH3_OK;
}

int ph3_host_free(void* ptr) {
    if (ptr) PH3_CHECK(hipHostFree(ptr));
    return PH3_OK;
}

static int copy_kind(int kind, hipMemcpyKind* k) {
    switch (kind) {
        case 0: *k = hipMemcpyHostToDevice; return PH3_OK;
        case 1: *k = hipMemcpyDeviceToHost; return PH3_OK;
        case 2: *k = hipMemcpyDeviceToDevice; return PH3_OK;
    }
    set_error("memcpy kind %d is not 0 (h2d), 1 (d2h) or 2 (d2d)", kind);
    return PH3_EINVAL;
}

// Complete before returning: hipMemcpy alone may return before a device-to-device copy (or the
// DMA of a pageable host-to-device copy) has landed, and kernels on non-blocking streams are not
// ordered after the null stream, so the copy is followed by a wait on the stream it ran on.
int ph3_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    hipMemcpyKind k;
    if (int r = copy_kind(kind, &k)) return r;
    if (bytes) {
        PH3_CHECK(hipMemcpyAsync(dst, src, bytes, k, nullptr));
        PH3_CHECK(hipStreamSynchronize(nullptr));
    }
    return PH3_OK;
}

int ph3_memcpy_async(void* dst, const void* src, size_t bytes, int kind, void* stream) {
    hipMemcpyKind k;
    if (int r = copy_kind(kind, &k)) return r;
    if (bytes) PH3_CHECK(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
    return PH3_OK;
}

int ph3_stream_create(void** stream) {
    hipStream_t s;
    PH3_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return PH3_OK;
}

int ph3_stream_destroy(void* stream) {
    if (stream) PH3_CHECK(hipStreamDestroy((hipStream_t)stream));
    return PH3_OK;
}

int ph3_stream_sync(void* stream) {
    PH3_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return PH3_OK;
}

int ph3_event_create(void** event) {
    hipEvent_t e;
    PH3_CHECK(hipEventCreate(&e));
    *event = (void*)e;
    return PH3_OK;
}

int ph3_event_destroy(void* event) {
    if (event) PH3_CHECK(hipEventDestroy((hipEvent_t)event));
    return PH3_OK;
}

int ph3_event_record(void* event, void* stream) {
    PH3_CHECK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return PH3_OK;
}

int ph3_event_sync(void* event) {
    PH3_CHECK(hipEventSynchronize((hipEvent_t)event));
    return PH3_OK;
}

int ph3_event_elapsed_ms(float* ms, void* start, void* stop) {
    PH3_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return PH3_OK;
}

int ph3_graph_begin(void* stream) {
    if (!stream) { set_error("graph capture needs a created stream, not the null stream"); return PH3_EINVAL; }
    PH3_CHECK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return PH3_OK;
}

int ph3_graph_end(void* stream, void** graph_exec) {
    if (!graph_exec) { set_error("ph3_graph_end: graph_exec is NULL"); return PH3_EINVAL; }
    *graph_exec = nullptr;
    hipGraph_t g = nullptr;
    PH3_CHECK(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t e = nullptr;
    const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (r != hipSuccess) return ph3host::check(r, "hipGraphInstantiate");
    *graph_exec = (void*)e;
    return PH3_OK;
}

int ph3_graph_launch(void* graph_exec, void* stream) {
    PH3_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return PH3_OK;
}

int ph3_graph_destroy(void* graph_exec) {
    if (graph_exec) PH3_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return PH3_OK;
}

}  // extern "C"
