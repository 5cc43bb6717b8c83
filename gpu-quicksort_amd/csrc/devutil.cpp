// devutil.cpp -- device memory, stream and event helpers of the C ABI.
//
// Callers that drive the library through a foreign-function layer (ctypes,
// cgo, JNI) need device buffers, streams and timers from the same HIP runtime
// the library links; these thin wrappers provide them.
#include <hip/hip_runtime.h>

#include "../../include/gqsort.h"
#include "internal.h"

extern "C" {

gqs_status gqs_malloc(gqs_ctx* c, size_t bytes, void** out) {
    if (!c || !out) return GQS_EINVAL;
    *out = nullptr;
    if (c->fail(hipSetDevice(c->device), "hipSetDevice")) return GQS_EDEVICE;
    if (c->fail(hipMalloc(out, bytes ? bytes : 1), "hipMalloc")) return GQS_ENOMEM;
    return GQS_OK;
}

gqs_status gqs_free(gqs_ctx* c, void* p) {
    if (!c) return GQS_EINVAL;
    if (c->fail(hipFree(p), "hipFree")) return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_host_alloc(gqs_ctx* c, size_t bytes, void** out) {
    if (!c || !out) return GQS_EINVAL;
    *out = nullptr;
    if (c->fail(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc")) return GQS_ENOMEM;
    return GQS_OK;
}

gqs_status gqs_host_free(gqs_ctx* c, void* p) {
    if (!c) return GQS_EINVAL;
    if (c->fail(hipHostFree(p), "hipHostFree")) return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_memcpy(gqs_ctx* c, void* dst, const void* src, size_t bytes, void* stream) {
    if (!c || (bytes && (!dst || !src))) return GQS_EINVAL;
    if (bytes == 0) return GQS_OK;
    hipStream_t st = (hipStream_t)stream;
    if (c->fail(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st), "hipMemcpyAsync") ||
        c->fail(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_memcpy_async(gqs_ctx* c, void* dst, const void* src, size_t bytes, void* stream) {
    if (!c || (bytes && (!dst || !src))) return GQS_EINVAL;
    if (bytes == 0) return GQS_OK;
    if (c->fail(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream), "hipMemcpyAsync"))
        return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_stream_create(gqs_ctx* c, void** out) {
    if (!c || !out) return GQS_EINVAL;
    hipStream_t s = nullptr;
    if (c->fail(hipSetDevice(c->device), "hipSetDevice") ||
        c->fail(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"))
        return GQS_EDEVICE;
    *out = (void*)s;
    return GQS_OK;
}

gqs_status gqs_stream_destroy(gqs_ctx* c, void* s) {
    if (!c) return GQS_EINVAL;
    if (c->fail(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy")) return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_stream_sync(gqs_ctx* c, void* s) {
    if (!c) return GQS_EINVAL;
    if (c->fail(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize")) return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_event_create(gqs_ctx* c, void** out) {
    if (!c || !out) return GQS_EINVAL;
    hipEvent_t e = nullptr;
    if (c->fail(hipEventCreate(&e), "hipEventCreate")) return GQS_EDEVICE;
    *out = (void*)e;
    return GQS_OK;
}

gqs_status gqs_event_destroy(gqs_ctx* c, void* e) {
    if (!c) return GQS_EINVAL;
    if (c->fail(hipEventDestroy((hipEvent_t)e), "hipEventDestroy")) return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_event_record(gqs_ctx* c, void* e, void* stream) {
    if (!c || !e) return GQS_EINVAL;
    if (c->fail(hipEventRecord((hipEvent_t)e, (hipStream_t)stream), "hipEventRecord")) return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status gqs_event_elapsed(gqs_ctx* c, void* start, void* stop, float* ms) {
    if (!c || !start || !stop || !ms) return GQS_EINVAL;
    if (c->fail(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize") ||
        c->fail(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime"))
        return GQS_EDEVICE;
    return GQS_OK;
}

}  // extern "C"
