// autotune.cpp -- parameter search for the sort, per (dtype, size class).
//
// The reference fixes QUICKSORT_BLOCK_SIZE, SORT_THRESHOLD and the work-group
// sizes per vendor at compile time (CL12/Quicksort.h:77-98; FSY/Quicksort.h:
// 116-137).  Here they are run-time parameters; gqs_autotune times sorts of
// synthetic keys on the device and keeps the fastest setting.  The search is a
// coordinate descent: starting from the current parameters, each parameter in
// turn takes every value of its candidate list while the others stay fixed,
// and the best value is kept before moving to the next parameter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gqsort.h"
#include "internal.h"

using namespace gqs;

namespace {

int size_class(uint64_t n) {
    int lg = 0;
    while ((2ull << lg) <= n) ++lg;
    return lg;
}

struct Field {
    const char* name;
    uint32_t gqs_params::*ptr;
};
const Field kFields[] = {
    {"leaf_max", &gqs_params::leaf_max},
    {"leaf_target", &gqs_params::leaf_target},
    {"oversample", &gqs_params::oversample},
    {"max_fanout", &gqs_params::max_fanout},
    {"max_levels", &gqs_params::max_levels},
    {"host_driven", &gqs_params::host_driven},
    {"part_wgs_per_cu", &gqs_params::part_wgs_per_cu},
    {"leaf_wgs_per_cu", &gqs_params::leaf_wgs_per_cu},
    {"sort_threshold", &gqs_params::sort_threshold},
    {"bitonic_cutoff", &gqs_params::bitonic_cutoff},
};

}  // namespace

extern "C" {

gqs_status gqs_autotune(gqs_ctx* c, int dtype, size_t n, int gen, uint32_t reps, gqs_params* best) {
    if (!c || n < 2 || dtype < GQS_U32 || dtype > GQS_F64) return GQS_EINVAL;
    if (dtype != GQS_U32 && gen != 7 && gen != 6) gen = 7;  // float generators
    if (reps == 0) reps = 3;
    const size_t kb = dtype == GQS_F64 ? 8 : 4;
    std::vector<unsigned char> host(n * kb);
    if (gqs_status s = gqs_generate(dtype, gen, 12345, host.data(), n)) return s;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->fail(hipSetDevice(c->device), "hipSetDevice")) return GQS_EDEVICE;
    void *pr = nullptr, *d = nullptr, *dn = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&pr, n * kb);
    if (!e) e = hipMalloc(&d, n * kb);
    if (!e) e = hipMalloc(&dn, n * kb);
    if (!e) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (!e) e = hipEventCreate(&e0);
    if (!e) e = hipEventCreate(&e1);
    if (!e) e = hipMemcpy(pr, host.data(), n * kb, hipMemcpyHostToDevice);
    gqs_status status = e ? GQS_ENOMEM : GQS_OK;
    const bool saved_use = c->use_tuned;
    const gqs_params saved = c->params;
    c->use_tuned = false;
    auto time_params = [&](const gqs_params& p, double* ms) -> gqs_status {
        c->params = p;
        std::vector<float> ts;
        for (uint32_t r = 0; r <= reps; ++r) {  // first run warms the workspace
            if (c->fail(hipMemcpyAsync(d, pr, n * kb, hipMemcpyDeviceToDevice, st), "restore")) return GQS_EDEVICE;
            hipEventRecord(e0, st);
            if (gqs_status s = c->run(c->ws, dtype, d, dn, n, nullptr, 0, nullptr, st, false)) return s;
            hipEventRecord(e1, st);
            float t = 0.f;
            if (c->fail(hipEventSynchronize(e1), "sync") || c->fail(hipEventElapsedTime(&t, e0, e1), "elapsed"))
                return GQS_EDEVICE;
            if (r) ts.push_back(t);
        }
        std::sort(ts.begin(), ts.end());
        *ms = ts[ts.size() / 2];
        return GQS_OK;
    };
    gqs_params cur = saved;
    cur.host_driven = 0;
    const uint32_t lmax = 2 * gqs::leaf_cap(dtype);
    cur.leaf_max = std::min(cur.leaf_max, lmax);
    double best_ms = 0;
    if (status == GQS_OK) status = time_params(cur, &best_ms);
    struct Axis {
        uint32_t gqs_params::*ptr;
        std::vector<uint32_t> values;
    };
    const std::vector<Axis> axes = {
        {&gqs_params::part_wgs_per_cu, {2, 4, 8}},
        {&gqs_params::leaf_wgs_per_cu, {2, 4, 8}},
        {&gqs_params::leaf_max, {lmax / 2, lmax}},
        {&gqs_params::leaf_target, {0, 0}},  // filled from leaf_max below
        {&gqs_params::max_fanout, {128, 256}},
        {&gqs_params::oversample, {8, 16, 32}},
        {&gqs_params::sort_threshold, {8, 16, 32}},
        {&gqs_params::bitonic_cutoff, {256, 512}},
    };
    for (const Axis& ax : axes) {
        if (status != GQS_OK) break;
        std::vector<uint32_t> vals = ax.values;
        // the fan-out is ceil(n / leaf_target): targets between a quarter and a
        // half of leaf_max trade leaves above one work-group's capacity
        // against half-empty ones
        if (ax.ptr == &gqs_params::leaf_target)
            vals = {cur.leaf_max / 4, cur.leaf_max * 5 / 16, cur.leaf_max * 3 / 8, cur.leaf_max * 7 / 16,
                    cur.leaf_max / 2};
        for (uint32_t v : vals) {
            if (cur.*ax.ptr == v) continue;
            gqs_params cand = cur;
            cand.*ax.ptr = v;
            if (cand.leaf_target > cand.leaf_max) continue;
            double ms = 0;
            status = time_params(cand, &ms);
            if (status != GQS_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                cur = cand;
            }
        }
    }
    c->params = saved;
    c->use_tuned = saved_use;
    if (status == GQS_OK) {
        const auto key = std::make_pair(dtype, size_class(n));
        c->tuned[key] = cur;
        c->tuned_ms[key] = best_ms;
        c->use_tuned = true;
        if (best) *best = cur;
    }
    if (st) hipStreamSynchronize(st);
    hipFree(pr);
    hipFree(d);
    hipFree(dn);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    return status;
}

gqs_status gqs_tune_save(gqs_ctx* c, const char* path) {
    if (!c || !path) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    std::ofstream f(path);
    if (!f) return c->set_error(GQS_EINVAL, "gqs_tune_save: cannot open file");
    f << "{\"device\": \"" << c->name << "\", \"entries\": [\n";
    bool first = true;
    for (const auto& kv : c->tuned) {
        const gqs_params& p = kv.second;
        f << (first ? "" : ",\n") << "  {\"dtype\": " << kv.first.first << ", \"log2n\": " << kv.first.second
          << ", \"ms\": " << c->tuned_ms[kv.first] << ", \"mode\": " << p.mode;
        for (const Field& fd : kFields) f << ", \"" << fd.name << "\": " << p.*fd.ptr;
        f << "}";
        first = false;
    }
    f << "\n]}\n";
    return f ? GQS_OK : c->set_error(GQS_EINVAL, "gqs_tune_save: write failed");
}

// Reads the format gqs_tune_save writes: one JSON object per entry line with
// integer fields (unknown fields keep their defaults).  Every entry must name
// a known key type and pass the range checks of gqs_set_params; otherwise the
// file is rejected as a whole and the context's table is left unchanged.
gqs_status gqs_tune_load(gqs_ctx* c, const char* path) {
    if (!c || !path) return GQS_EINVAL;
    std::ifstream f(path);
    if (!f) return c->set_error(GQS_EINVAL, "gqs_tune_load: cannot open file");
    std::lock_guard<std::mutex> lk(c->mu);
    std::string line;
    auto field = [](const std::string& l, const std::string& name, double* out) {
        const std::string key = "\"" + name + "\":";
        const size_t at = l.find(key);
        if (at == std::string::npos) return false;
        *out = std::strtod(l.c_str() + at + key.size(), nullptr);
        return true;
    };
    std::map<std::pair<int, int>, gqs_params> table;
    std::map<std::pair<int, int>, double> table_ms;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        double dt, lg, v;
        if (!field(line, "dtype", &dt) || !field(line, "log2n", &lg)) continue;
        char msg[128];
        if (dt != GQS_U32 && dt != GQS_F32 && dt != GQS_F64) {
            std::snprintf(msg, sizeof msg, "gqs_tune_load: line %d: unknown dtype", lineno);
            return c->set_error(GQS_EINVAL, msg);
        }
        if (lg < 1 || lg > 63) {
            std::snprintf(msg, sizeof msg, "gqs_tune_load: line %d: log2n out of range", lineno);
            return c->set_error(GQS_EINVAL, msg);
        }
        gqs_params p{};
        gqs_default_params((int)dt, &p);
        if (field(line, "mode", &v)) p.mode = (int32_t)v;
        bool in_range = true;
        for (const Field& fd : kFields)
            if (field(line, fd.name, &v)) {
                in_range = in_range && v >= 0 && v <= 4294967295.0;
                p.*fd.ptr = (uint32_t)v;
            }
        if (!in_range || !gqs::params_valid(p)) {
            std::snprintf(msg, sizeof msg, "gqs_tune_load: line %d: parameter out of range", lineno);
            return c->set_error(GQS_EINVAL, msg);
        }
        const auto key = std::make_pair((int)dt, (int)lg);
        table[key] = p;
        if (field(line, "ms", &v)) table_ms[key] = v;
    }
    if (table.empty()) return c->set_error(GQS_EINVAL, "gqs_tune_load: no entries");
    for (const auto& kv : table) c->tuned[kv.first] = kv.second;
    for (const auto& kv : table_ms) c->tuned_ms[kv.first] = kv.second;
    c->use_tuned = true;
    return GQS_OK;
}

}  // extern "C"
