// serve.cpp -- continuous-batching sort server over the segmented sort.
//
// The reference has no serving path; its only caller sorts one array per
// GPUQSort call (CL12/QuicksortMain.cpp:485-497).  This server admits sort
// requests of mixed sizes and dtypes while earlier batches are still on the
// device:
//   * `slots` independent pipelines, each with its own HIP stream, its own
//     sort workspace and its own device-resident key buffers (d, dn) of
//     `slot_keys` keys, all allocated at creation inside the HBM budget;
//   * a scheduler thread that, whenever a slot is free, packs the oldest
//     pending requests of one dtype into a batch (up to the slot's capacity),
//     copies them in, runs one segmented sort, copies the results out and
//     records an event; finished batches are retired by polling their events,
//     so new requests are admitted as soon as a slot frees up;
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

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    void* d = nullptr;
    void* dn = nullptr;
    uint64_t* offs_dev = nullptr;
    uint64_t* offs_host = nullptr;  // pinned
    std::vector<Request*> batch;
    size_t batch_keys = 0;
    bool busy = false;
};

}  // namespace

struct gqs_server {
    gqs_ctx* ctx = nullptr;
    gqs_server_config cfg{};
    std::vector<Slot> slots;
    std::mutex m;
    std::condition_variable cv_submit, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    std::atomic<bool> stop{false};
    std::thread sched;
    gqs_server_stats stats{};
    std::string err;

    void retire();
    bool admit();
    void loop();
};

void gqs_server::retire() {
    for (auto& s : slots) {
        if (!s.busy) continue;
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipErrorNotReady) continue;
        const double t = now_us();
        std::lock_guard<std::mutex> lk(m);
        for (Request* r : s.batch) {
            r->t_done = t;
            r->state = (q == hipSuccess) ? 2 : 3;
        }
        stats.completed += s.batch.size();
        stats.keys_sorted += s.batch_keys;
        s.batch.clear();
        s.batch_keys = 0;
        s.busy = false;
        cv_done.notify_all();
    }
}

// Packs pending requests into the first free slot.  Returns true if a batch
// was launched.
bool gqs_server::admit() {
    Slot* slot = nullptr;
    for (auto& s : slots)
        if (!s.busy) {
            slot = &s;
            break;
        }
    if (!slot) return false;
    std::vector<Request*> batch;
    size_t keys = 0;
    int dtype = -1;
    {
        std::lock_guard<std::mutex> lk(m);
        if (pending.empty()) return false;
        dtype = pending.front()->dtype;
        const size_t kb = dtype == GQS_F64 ? 8 : 4;
        const size_t cap_keys = cfg.slot_keys * 8 / kb;  // slot buffers are sized in 8-byte keys
        for (auto it = pending.begin(); it != pending.end() && batch.size() < cfg.max_batch_requests;) {
            Request* r = *it;
            if (r->dtype == dtype && keys + r->n <= cap_keys) {
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
    for (size_t i = 0; i < batch.size() && !e; ++i)
        e = hipMemcpyAsync(static_cast<char*>(slot->d) + slot->offs_host[i] * kb, batch[i]->in, batch[i]->n * kb,
                           hipMemcpyHostToDevice, slot->stream);
    if (!e)
        e = hipMemcpyAsync(slot->offs_dev, slot->offs_host, (batch.size() + 1) * sizeof(uint64_t),
                           hipMemcpyHostToDevice, slot->stream);
    gqs_status st = GQS_OK;
    if (!e) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        st = ctx->run(slot->ws, dtype, slot->d, slot->dn, off, slot->offs_dev, (uint32_t)batch.size(), nullptr,
                      slot->stream, false);
    }
    for (size_t i = 0; i < batch.size() && !e && st == GQS_OK; ++i)
        e = hipMemcpyAsync(batch[i]->out, static_cast<char*>(slot->d) + slot->offs_host[i] * kb, batch[i]->n * kb,
                           hipMemcpyDeviceToHost, slot->stream);
    if (!e) e = hipEventRecord(slot->done, slot->stream);
    if (e || st != GQS_OK) {
        std::lock_guard<std::mutex> lk(m);
        for (Request* r : batch) {
            r->state = 3;
            r->t_done = now_us();
        }
        stats.failed += batch.size();
        cv_done.notify_all();
        return false;
    }
    slot->batch = std::move(batch);
    slot->batch_keys = off;
    slot->busy = true;
    std::lock_guard<std::mutex> lk(m);
    stats.batches++;
    stats.max_batch_requests = std::max<uint64_t>(stats.max_batch_requests, slot->batch.size());
    return true;
}

void gqs_server::loop() {
    hipSetDevice(ctx->device);
    while (!stop.load()) {
        retire();
        bool launched = false;
        while (admit()) launched = true;
        if (launched) continue;
        bool any_busy = false;
        for (auto& s : slots) any_busy |= s.busy;
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
    c->slots = 3;
    c->slot_keys = 0;  // derived from the budget
    c->max_batch_requests = 4096;
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
    for (auto& sl : s->slots) {
        if (!e) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (!e) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
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
        for (auto& sl : s->slots) {
            hipFree(sl.d);
            hipFree(sl.dn);
            hipFree(sl.offs_dev);
            hipHostFree(sl.offs_host);
            sl.ws.release_all();
            if (sl.done) hipEventDestroy(sl.done);
            if (sl.stream) hipStreamDestroy(sl.stream);
        }
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
    for (auto& sl : s->slots) {
        if (sl.stream) hipStreamSynchronize(sl.stream);
        hipFree(sl.d);
        hipFree(sl.dn);
        hipFree(sl.offs_dev);
        hipHostFree(sl.offs_host);
        sl.ws.release_all();
        if (sl.done) hipEventDestroy(sl.done);
        if (sl.stream) hipStreamDestroy(sl.stream);
    }
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
