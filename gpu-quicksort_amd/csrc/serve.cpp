// serve.cpp -- continuous-batching sort server over the segmented sort.
//
// The reference has no serving path; its only caller sorts one array per
// GPUQSort call (CL12/QuicksortMain.cpp:485-497).  This server admits sort
// requests of mixed sizes and dtypes while earlier batches are still on the
// device, as a three-stage pipeline:
//   * `slots` batches in flight, each slot with its own sort workspace and
//     device-resident key buffers (d, dn) of `slot_keys` keys, all allocated
//     at creation inside the HBM budget;
//   * three streams shared by the slots, one per stage (upload, sort,
//     download): batch k+1's upload, batch k's sort and batch k-1's download
//     run at once, so both directions of the PCIe link and the GPU are busy
//     together.  The scheduler moves a batch to its next stage when it sees
//     (by polling the stage's event) that the previous one has finished, so
//     no stream waits on another stream's event: on this runtime a copy
//     enqueued behind such a wait blocks the enqueuing thread until the wait
//     is satisfied, which would stall every other batch's progress too;
//   * a scheduler thread that, whenever a slot is free, packs the oldest
//     pending requests of one dtype into a batch of about `batch_bytes`
//     (more requests wait less behind one copy; a larger request gets a
//     batch of its own), and retires requests one by one as their own
//     download events fire, so a request is done when its keys are back,
//     not when its whole batch is;
//   * per-request timestamps (submit, admit, done) for latency statistics.
// A request larger than a slot is rejected at submit time (GQS_EBUSY).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gqsort.h"
#include "internal.h"

using namespace gqs;

namespace {

constexpr uint64_t kDefaultBatchBytes = 64ull << 20;  // batch target: about 1 ms of PCIe each way
constexpr int kUploadsAhead = 2;  // batches queued on the upload stream before another is formed
constexpr size_t kCopyChunk = 32ull << 20;  // largest single DMA of a request's keys

// A copy as DMAs of at most kCopyChunk bytes: a large request does not hold
// a copy engine in one piece while the other direction's transfers wait.
hipError_t chunked_copy(char* dst, const char* src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
    hipError_t e = hipSuccess;
    for (size_t o = 0; o < bytes && e == hipSuccess; o += kCopyChunk)
        e = hipMemcpyAsync(dst + o, src + o, std::min(kCopyChunk, bytes - o), kind, st);
    return e;
}

double now_us() {
    using namespace std::chrono;
    return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    int dtype = 0;
    const void* in = nullptr;
    void* out = nullptr;
    size_t n = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = 0;  // 0 pending, 1 running, 2 done, 3 failed
};

// A slot's batch goes through the stages in order; `stage` is the one it is
// in (its work enqueued, its event not yet seen).
enum Stage : int { FREE = 0, UPLOADING, SORTING, DOWNLOADING };

struct Slot {
    hipEvent_t uploaded = nullptr;  // keys and offsets on the device (upload stream)
    hipEvent_t sorted = nullptr;    // the sort is done (sort stream)
    Workspace ws;
    void* d = nullptr;
    void* dn = nullptr;
    uint64_t* offs_dev = nullptr;
    uint64_t* offs_host = nullptr;  // pinned
    std::vector<Request*> batch;
    std::vector<hipEvent_t> req_done;  // per request: its keys are back (download stream), grown on demand
    size_t retired = 0;                // requests of the batch already marked done
    size_t batch_keys = 0;
    int dtype = 0;
    uint64_t seq = 0;  // admission order
    Stage stage = FREE;
};

}  // namespace

struct gqs_server {
    gqs_ctx* ctx = nullptr;
    gqs_server_config cfg{};
    std::vector<Slot> slots;
    hipStream_t up = nullptr, sorter = nullptr, down = nullptr;  // the three stages, each FIFO
    std::mutex m;
    std::condition_variable cv_submit, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    std::atomic<bool> stop{false};
    std::thread sched;
    gqs_server_stats stats{};
    std::string err;

    uint64_t next_seq = 0;

    bool advance();
    bool admit();
    void loop();
    void release_device();
    void fail_batch(Slot& slot);
};

// Waits for and frees every stream, event and buffer of the server.
void gqs_server::release_device() {
    for (hipStream_t st : {up, sorter, down})
        if (st) hipStreamSynchronize(st);
    for (auto& sl : slots) {
        hipFree(sl.d);
        hipFree(sl.dn);
        hipFree(sl.offs_dev);
        hipHostFree(sl.offs_host);
        sl.ws.release_all();
        for (hipEvent_t ev : {sl.uploaded, sl.sorted})
            if (ev) hipEventDestroy(ev);
        for (hipEvent_t ev : sl.req_done) hipEventDestroy(ev);
        sl.req_done.clear();
        sl.uploaded = sl.sorted = nullptr;
    }
    for (hipStream_t st : {up, sorter, down})
        if (st) hipStreamDestroy(st);
    up = sorter = down = nullptr;
}

// Moves every batch whose current stage has finished on to the next one, in
// admission order (each stage's stream is FIFO): uploaded -> enqueue its
// sort; sorted -> enqueue its downloads, one event per request; each
// request is done as its own download event fires, and the slot is free
// after the last.  Returns true if anything moved.
bool gqs_server::advance() {
    std::vector<Slot*> order;
    for (auto& s : slots)
        if (s.stage != FREE) order.push_back(&s);
    std::sort(order.begin(), order.end(), [](const Slot* a, const Slot* b) { return a->seq < b->seq; });
    bool moved = false;
    for (Slot* sp : order) {
        Slot& s = *sp;
        if (s.stage == UPLOADING) {
            const hipError_t q = hipEventQuery(s.uploaded);
            if (q == hipErrorNotReady) continue;
            gqs_status st = q == hipSuccess ? GQS_OK : GQS_EDEVICE;
            if (st == GQS_OK) {
                std::lock_guard<std::mutex> lk(ctx->mu);
                st = ctx->run(s.ws, s.dtype, s.d, s.dn, s.batch_keys, s.offs_dev, (uint32_t)s.batch.size(), nullptr,
                              sorter, false);
            }
            if (st != GQS_OK || hipEventRecord(s.sorted, sorter) != hipSuccess) {
                fail_batch(s);
                continue;
            }
            s.stage = SORTING;
            moved = true;
        }
        if (s.stage == SORTING) {
            const hipError_t q = hipEventQuery(s.sorted);
            if (q == hipErrorNotReady) continue;
            const size_t kb = s.dtype == GQS_F64 ? 8 : 4;
            hipError_t e = q;
            for (size_t i = 0; i < s.batch.size() && e == hipSuccess; ++i) {
                e = chunked_copy(static_cast<char*>(s.batch[i]->out),
                                 static_cast<const char*>(s.d) + s.offs_host[i] * kb, s.batch[i]->n * kb,
                                 hipMemcpyDeviceToHost, down);
                if (e == hipSuccess) e = hipEventRecord(s.req_done[i], down);
            }
            if (e != hipSuccess) {
                fail_batch(s);
                continue;
            }
            s.stage = DOWNLOADING;
            moved = true;
        }
        if (s.stage == DOWNLOADING) {
            bool any = false;
            while (s.retired < s.batch.size()) {
                const hipError_t q = hipEventQuery(s.req_done[s.retired]);
                if (q == hipErrorNotReady) break;
                const double t = now_us();
                std::lock_guard<std::mutex> lk(m);
                Request* r = s.batch[s.retired++];
                r->t_done = t;
                r->state = (q == hipSuccess) ? 2 : 3;
                if (q == hipSuccess) stats.completed++;
                else stats.failed++;
                any = true;
            }
            if (s.retired == s.batch.size()) {
                std::lock_guard<std::mutex> lk(m);
                stats.keys_sorted += s.batch_keys;
                s.batch.clear();
                s.retired = 0;
                s.batch_keys = 0;
                s.stage = FREE;
            }
            if (any) {
                cv_done.notify_all();
                moved = true;
            }
        }
    }
    return moved;
}

// Marks the requests of a batch that have not finished as failed and frees
// its slot once the device is done with its buffers.
void gqs_server::fail_batch(Slot& slot) {
    for (hipStream_t st : {up, sorter, down}) hipStreamSynchronize(st);
    std::lock_guard<std::mutex> lk(m);
    for (size_t i = slot.retired; i < slot.batch.size(); ++i) {
        slot.batch[i]->state = 3;
        slot.batch[i]->t_done = now_us();
        stats.failed++;
    }
    slot.batch.clear();
    slot.retired = 0;
    slot.batch_keys = 0;
    slot.stage = FREE;
    cv_done.notify_all();
}

// Packs pending requests into the first free slot.  Returns true if a batch
// was launched.  A batch is formed only while fewer than kUploadsAhead
// batches wait for or occupy the upload stream: the link stays busy, and
// requests that arrive meanwhile join the next batch instead of each paying
// a sort launch of its own.
bool gqs_server::admit() {
    Slot* slot = nullptr;
    int uploading = 0;
    for (auto& s : slots) {
        if (s.stage == FREE) {
            if (!slot) slot = &s;
        } else if (s.stage == UPLOADING) {
            ++uploading;
        }
    }
    if (!slot || uploading >= kUploadsAhead) return false;
    std::vector<Request*> batch;
    size_t keys = 0;
    int dtype = -1;
    {
        std::lock_guard<std::mutex> lk(m);
        if (pending.empty()) return false;
        dtype = pending.front()->dtype;
        const size_t kb = dtype == GQS_F64 ? 8 : 4;
        const size_t cap_keys = cfg.slot_keys * 8 / kb;  // slot buffers are sized in 8-byte keys
        // the oldest request always goes in; others while the batch stays
        // within batch_bytes (smaller ones may pass a larger one)
        const size_t target_keys = std::max<size_t>(1, cfg.batch_bytes / kb);
        for (auto it = pending.begin(); it != pending.end() && batch.size() < cfg.max_batch_requests;) {
            Request* r = *it;
            if (r->dtype == dtype && keys + r->n <= cap_keys && (batch.empty() || keys + r->n <= target_keys)) {
                batch.push_back(r);
                keys += r->n;
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        const double t = now_us();
        for (Request* r : batch) {
            r->t_admit = t;
            r->state = 1;
        }
    }
    if (batch.empty()) return false;
    const size_t kb = dtype == GQS_F64 ? 8 : 4;
    uint64_t off = 0;
    for (size_t i = 0; i < batch.size(); ++i) {
        slot->offs_host[i] = off;
        off += batch[i]->n;
    }
    slot->offs_host[batch.size()] = off;
    hipError_t e = hipSuccess;
    while (!e && slot->req_done.size() < batch.size()) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (!e) slot->req_done.push_back(ev);
    }
    // upload (shared FIFO stream), in DMA pieces of at most kCopyChunk bytes
    for (size_t i = 0; i < batch.size() && !e; ++i)
        e = chunked_copy(static_cast<char*>(slot->d) + slot->offs_host[i] * kb, static_cast<const char*>(batch[i]->in),
                         batch[i]->n * kb, hipMemcpyHostToDevice, up);
    if (!e)
        e = hipMemcpyAsync(slot->offs_dev, slot->offs_host, (batch.size() + 1) * sizeof(uint64_t),
                           hipMemcpyHostToDevice, up);
    if (!e) e = hipEventRecord(slot->uploaded, up);
    if (e) {
        slot->batch = std::move(batch);
        fail_batch(*slot);
        return false;
    }
    slot->batch = std::move(batch);
    slot->retired = 0;
    slot->batch_keys = off;
    slot->dtype = dtype;
    slot->seq = next_seq++;
    slot->stage = UPLOADING;
    std::lock_guard<std::mutex> lk(m);
    stats.batches++;
    stats.max_batch_requests = std::max<uint64_t>(stats.max_batch_requests, slot->batch.size());
    return true;
}

void gqs_server::loop() {
    hipSetDevice(ctx->device);
    while (!stop.load()) {
        bool moved = advance();
        while (admit()) moved = true;
        if (moved) continue;
        bool any_busy = false;
        for (auto& s : slots) any_busy |= s.stage != FREE;
        std::unique_lock<std::mutex> lk(m);
        if (stop.load()) break;
        if (any_busy) {
            // Poll in-flight batches at a fine grain; a new submission wakes us too.
            cv_submit.wait_for(lk, std::chrono::microseconds(20));
        } else if (pending.empty()) {
            cv_submit.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

extern "C" {

gqs_status gqs_server_default_config(gqs_server_config* c) {
    if (!c) return GQS_EINVAL;
    c->hbm_budget_bytes = 64ull << 30;
    c->slots = 8;  // 2 uploading + 1 sorting + the downloads queued behind them
    c->slot_keys = 0;  // derived from the budget
    c->max_batch_requests = 4096;
    c->batch_bytes = 0;  // the default target
    return GQS_OK;
}

// Defaults and derived values of a server configuration.  A slot holds
// slot_keys 8-byte keys, i.e. twice as many 4-byte keys; one segmented sort
// takes fewer than 2^32 - 2^16 keys, so slot_keys is clamped to keep every
// batch of 4-byte keys below that.
gqs_status gqs_server_resolve_config(const gqs_server_config* in, gqs_server_config* out) {
    if (!out) return GQS_EINVAL;
    gqs_server_config c{};
    gqs_server_default_config(&c);
    if (in) c = *in;
    if (c.slots < 1 || c.slots > 16 || c.max_batch_requests < 1 || c.hbm_budget_bytes == 0) return GQS_EINVAL;
    // Budget: per slot, d and dn of slot_keys 8-byte keys plus the sort
    // workspace (sized for slot_keys keys and as many arrays).
    if (c.slot_keys == 0) c.slot_keys = c.hbm_budget_bytes / c.slots / 24;
    c.slot_keys = std::min<uint64_t>(c.slot_keys, (gqs::kMaxRunKeys - 1) / 2);
    if (c.slot_keys < 1) return GQS_EINVAL;
    if (c.batch_bytes == 0) c.batch_bytes = kDefaultBatchBytes;
    *out = c;
    return GQS_OK;
}

gqs_status gqs_server_create(int device, const gqs_server_config* cfg, gqs_server** out) {
    if (!out) return GQS_EINVAL;
    *out = nullptr;
    gqs_server_config c{};
    if (gqs_status st = gqs_server_resolve_config(cfg, &c)) return st;
    auto* s = new (std::nothrow) gqs_server();
    if (!s) return GQS_ENOMEM;
    if (gqs_status st = gqs_create(&s->ctx, device)) {
        delete s;
        return st;
    }
    s->cfg = c;
    s->slots.resize(c.slots);
    hipError_t e = hipSetDevice(device);
    if (!e) e = hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking);
    if (!e) e = hipStreamCreateWithFlags(&s->sorter, hipStreamNonBlocking);
    if (!e) e = hipStreamCreateWithFlags(&s->down, hipStreamNonBlocking);
    for (auto& sl : s->slots) {
        if (!e) e = hipEventCreateWithFlags(&sl.uploaded, hipEventDisableTiming);
        if (!e) e = hipEventCreateWithFlags(&sl.sorted, hipEventDisableTiming);
        if (!e) e = hipMalloc(&sl.d, c.slot_keys * 8);
        if (!e) e = hipMalloc(&sl.dn, c.slot_keys * 8);
        if (!e) e = hipMalloc(reinterpret_cast<void**>(&sl.offs_dev), (c.max_batch_requests + 1) * 8);
        if (!e) e = hipHostMalloc(reinterpret_cast<void**>(&sl.offs_host), (c.max_batch_requests + 1) * 8,
                                  hipHostMallocDefault);
        if (e) break;
        // one workspace per slot, big enough for a batch of any key type and
        // the parameters of any size class it may run with
        for (int dt : {GQS_U32, GQS_F32, GQS_F64}) {
            const uint64_t keys = dt == GQS_F64 ? c.slot_keys : c.slot_keys * 2;
            if (s->ctx->ensure_caps(sl.ws, s->ctx->caps_upto(keys, c.max_batch_requests, dt), dt) != GQS_OK)
                e = hipErrorOutOfMemory;
        }
        if (e) break;
    }
    if (e) {
        s->release_device();
        gqs_destroy(s->ctx);
        delete s;
        return GQS_ENOMEM;
    }
    s->sched = std::thread([s] { s->loop(); });
    *out = s;
    return GQS_OK;
}

gqs_status gqs_server_submit(gqs_server* s, int dtype, const void* in, void* out, size_t n, uint64_t* ticket) {
    if (!s || !ticket || (n && (!in || !out)) || dtype < GQS_U32 || dtype > GQS_F64) return GQS_EINVAL;
    const size_t kb = dtype == GQS_F64 ? 8 : 4;
    // never fits a slot (slots hold fewer keys than one segmented sort's limit)
    if (n * kb > s->cfg.slot_keys * 8 || n >= gqs::kMaxRunKeys) return GQS_EBUSY;
    auto* r = new (std::nothrow) Request();
    if (!r) return GQS_ENOMEM;
    r->dtype = dtype;
    r->in = in;
    r->out = out;
    r->n = n;
    r->t_submit = now_us();
    {
        std::lock_guard<std::mutex> lk(s->m);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        if (n < 2) {  // nothing to sort: copy through
            if (n == 1 && out != in) std::memcpy(out, in, kb);
            r->state = 2;
            r->t_admit = r->t_done = r->t_submit;
            s->stats.completed++;
        } else {
            s->pending.push_back(r);
        }
        s->stats.submitted++;
    }
    s->cv_submit.notify_one();
    *ticket = r->id;
    return GQS_OK;
}

gqs_status gqs_server_wait(gqs_server* s, uint64_t ticket, int64_t timeout_us) {
    if (!s) return GQS_EINVAL;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GQS_ENOTFOUND;
    Request* r = it->second;
    auto pred = [r] { return r->state >= 2; };
    if (timeout_us < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), pred)) return GQS_EBUSY;
    return r->state == 2 ? GQS_OK : GQS_EDEVICE;
}

gqs_status gqs_server_poll(gqs_server* s, uint64_t ticket, int* done) {
    if (!s || !done) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GQS_ENOTFOUND;
    *done = it->second->state >= 2 ? 1 : 0;
    return it->second->state == 3 ? GQS_EDEVICE : GQS_OK;
}

gqs_status gqs_server_request_times(gqs_server* s, uint64_t ticket, double* submit_us, double* admit_us,
                                    double* done_us) {
    if (!s) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GQS_ENOTFOUND;
    if (submit_us) *submit_us = it->second->t_submit;
    if (admit_us) *admit_us = it->second->t_admit;
    if (done_us) *done_us = it->second->t_done;
    return GQS_OK;
}

gqs_status gqs_server_release(gqs_server* s, uint64_t ticket) {
    if (!s) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GQS_ENOTFOUND;
    if (it->second->state < 2) return GQS_EBUSY;
    delete it->second;
    s->reqs.erase(it);
    return GQS_OK;
}

gqs_status gqs_server_get_config(const gqs_server* s, gqs_server_config* out) {
    if (!s || !out) return GQS_EINVAL;
    *out = s->cfg;
    return GQS_OK;
}

gqs_status gqs_server_get_stats(gqs_server* s, gqs_server_stats* out) {
    if (!s || !out) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->stats;
    out->pending = s->pending.size();
    return GQS_OK;
}

gqs_status gqs_server_destroy(gqs_server* s) {
    if (!s) return GQS_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop.store(true);
    }
    s->cv_submit.notify_all();
    if (s->sched.joinable()) s->sched.join();
    hipSetDevice(s->ctx->device);
    s->release_device();
    {
        std::lock_guard<std::mutex> lk(s->m);
        for (auto& kv : s->reqs) {
            if (kv.second->state < 2) kv.second->state = 3;
            delete kv.second;
        }
        s->reqs.clear();
        s->cv_done.notify_all();
    }
    gqs_destroy(s->ctx);
    delete s;
    return GQS_OK;
}

}  // extern "C"
