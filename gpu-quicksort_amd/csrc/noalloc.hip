// noalloc.hip -- gqs_sort_device_noalloc: the device-resident sort without
// any allocation, for callers that cannot call hipMalloc (code under graph
// capture, tight serving loops, harnesses that forbid it).
//
// The sort's workspace (segment queues, splitter tables, count matrix, bucket
// and fill records, control blocks) is carved out of `scratch`, a file-scope
// device array that every GPU the library runs on holds from the loading of
// the code object until the process ends.  Its device address is looked up
// once per device, at context creation (gqs_create, with the device current);
// host code only ever uses that address, never the array's name.  All calls
// on one GPU share the array: the level counters are cleared by the chain's
// first kernel on every call, and calls that overlap on one GPU (different
// contexts, streams or threads) race on it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/gqsort.h"
#include "internal.h"

namespace gqs {
namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of scratch, per device

// Carves a workspace with capacities `c` out of the scratch at `base`;
// returns the bytes it spans (256-byte aligned pieces).
size_t carve(Workspace& w, unsigned char* base, const Caps& c) {
    size_t off = 0;
    auto take = [&](auto** p, size_t count, size_t elem) {
        *p = reinterpret_cast<std::remove_reference_t<decltype(**p)>*>(base + off);
        off += (std::max<size_t>(1, count) * elem + 255) / 256 * 256;
    };
    take(&w.segs[0], c.seg, sizeof(SegRec));
    take(&w.segs[1], c.seg, sizeof(SegRec));
    take(&w.gather, c.seg, sizeof(SegGather));
    take(&w.plans, c.seg, sizeof(SegPlan));
    unsigned char* spl = nullptr;
    take(&spl, c.seg * kSplStride, 8);
    w.spl = spl;
    take(&w.meta, c.seg * kSplStride, sizeof(uint16_t));
    take(&w.lut, c.seg * kLutCells, sizeof(uint8_t));
    take(&w.mat, c.mat * kMatRegions, sizeof(uint32_t));
    take(&w.bsum, c.seg * kSplStride, sizeof(uint32_t));
    take(&w.part, c.part, sizeof(uint32_t));
    take(&w.bkt, c.bkt, sizeof(LeafRec));
    take(&w.fill, c.fill, sizeof(FillTask));
    take(&w.ctl, kCtlBlocks, sizeof(LevelCtl));
    take(&w.leaves, c.leaves, sizeof(LeafRec));
    take(&w.huge, c.huge, sizeof(LeafRec));
    w.cap = c;
    w.key_bytes = 8;
    return off;
}

}  // namespace

gqs_status noalloc_setup(int device, unsigned char** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return GQS_EDEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GQS_EDEVICE;
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[device];
    return GQS_OK;
}

}  // namespace gqs

extern "C" gqs_status gqs_sort_device_noalloc(gqs_ctx* c, int dtype, void* d, void* dn, size_t n, void* stream) {
    using namespace gqs;
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (dtype < DT_U32 || dtype > DT_F64) return c->set_error(GQS_EINVAL, "gqs_sort_device_noalloc: bad dtype");
    if (n < 2) return GQS_OK;
    if (!d || !dn) return c->set_error(GQS_EINVAL, "gqs_sort_device_noalloc: null buffer");
    if (n > kNoAllocMaxKeys) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "gqs_sort_device_noalloc: %zu keys, above the maximum of %llu", n,
                      (unsigned long long)kNoAllocMaxKeys);
        return c->set_error(GQS_EINVAL, msg);
    }
    if (!c->noalloc_scratch)
        return c->set_error(GQS_EDEVICE, "gqs_sort_device_noalloc: scratch address not found at context creation");
    if (c->fail(hipSetDevice(c->device), "hipSetDevice")) return GQS_EDEVICE;
    // capacities for the parameters this sort runs with; the chain's own
    // workspace check then passes without allocating
    const Caps need = caps_for(n, 0, c->run_params(dtype, n), dtype);
    Workspace w;
    const size_t bytes = carve(w, c->noalloc_scratch, need);
    if (bytes > kNoAllocScratchBytes) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "gqs_sort_device_noalloc: the workspace for %zu keys under the current parameters needs %zu "
                      "bytes, above the %zu-byte scratch",
                      n, bytes, kNoAllocScratchBytes);
        return c->set_error(GQS_EINVAL, msg);
    }
    // per-launch timing events are not created on this path
    const bool prof = c->profiling;
    c->profiling = false;
    const gqs_status s = c->run(w, dtype, d, dn, n, nullptr, 0, nullptr, (hipStream_t)stream, false);
    c->profiling = prof;
    return s;
}
