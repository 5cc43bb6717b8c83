// gqsort.cpp -- host dispatcher and C ABI of the GPU-Quicksort library.
//
// GPUQSort's host side (CL12/QuicksortMain.cpp:297-384) builds per-level
// block/parent records, launches gqsort_kernel, reads the results back and
// splits them into work/done, once per level, then launches lqsort_kernel.
// Here the whole schedule is enqueued up front: each level's bookkeeping runs
// on the device (kernels.hip), so a sort of any size is one stream-ordered
// chain with no host round trip.  `host_driven` restores the reference's
// loop shape (one readback per level, stop when the queue is empty).  The
// classic single-pivot mode's depth is data dependent (about 30 levels at
// 2^28 keys on random input, up to n/2 on an organ pipe); its median-of-3
// levels are followed by multipivot levels sized for the worst case, so it
// needs no readback either.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <thread>
#include <atomic>
#include <iterator>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gqsort.h"
#include "internal.h"
#include "kernels.h"

using namespace gqs;

namespace {

size_t key_bytes(int dtype) { return dtype == DT_F64 ? 8 : 4; }

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

}  // namespace

namespace gqs {

bool Ctx::fail(hipError_t e, const char* what) {
    if (e == hipSuccess) return false;
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    err = buf;
    return true;
}

gqs_status Ctx::set_error(gqs_status s, const char* msg) {
    err = msg;
    return s;
}

void Workspace::release() {
    hipFree(huge);
    hipFree(segs[0]);
    hipFree(segs[1]);
    hipFree(gather);
    hipFree(plans);
    hipFree(spl);
    hipFree(meta);
    hipFree(lut);
    hipFree(mat);
    hipFree(bsum);
    hipFree(part);
    hipFree(bkt);
    hipFree(fill);
    hipFree(ctl);
    hipFree(leaves);
    hipEvent_t ev = done;
    hipStream_t st = last;
    *this = Workspace{};
    done = ev;
    last = st;
}

void Workspace::release_all() {
    release();
    if (done) hipEventDestroy(done);
    done = nullptr;
    last = nullptr;
}

gqs_status Ctx::stream_enter(Workspace& w, hipStream_t st) {
    if (w.done && w.last != st && fail(hipStreamWaitEvent(st, w.done, 0), "stream order")) return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status Ctx::stream_leave(Workspace& w, hipStream_t st) {
    if (!w.done && fail(hipEventCreateWithFlags(&w.done, hipEventDisableTiming), "stream-order event"))
        return GQS_EDEVICE;
    if (fail(hipEventRecord(w.done, st), "stream-order event")) return GQS_EDEVICE;
    w.last = st;
    return GQS_OK;
}

// Capacities from provable bounds (n = total keys, S = segments per level):
//   every segment of a level holds > leaf_max keys  =>  S <= n/(leaf_max+1)
//   multipivot fan-out k <= max(2, 2*ceil(n_s/leaf_target)), buckets <= 2k-1
//   count matrix <= sum_s ntiles_s * (nbuckets_s + 1)   (+1: a tile-local
//                   level's row of tile sizes)
//               <= 257*(n/tile) + sum_s nbuckets_s + S
Caps caps_for(uint64_t n, uint32_t nseg_in, const gqs_params& p, int dtype) {
    Caps c;
    const uint64_t tile = (dtype == DT_F64) ? TileShape<uint64_t>::kTileKeys : TileShape<uint32_t>::kTileKeys;
    // (at least one segment per bucket of one segment: the second of two
    // tile-local levels takes every bucket of the first as a segment)
    c.seg = std::max<uint64_t>(n / ((uint64_t)p.leaf_max + 1) + 2, kSplStride + 2);
    // classic mode: single-pivot levels (3 buckets per segment), then
    // multipivot rescue levels with the rescue leaf target
    const uint64_t lt = std::max<uint64_t>(1, p.mode == MODE_CLASSIC ? rescue_leaf_target(p) : p.leaf_target);
    const uint64_t sum_nb = std::max<uint64_t>(3 * c.seg, 8 * c.seg + 4 * (n / lt + 1));
    c.bkt = sum_nb + 256;
    // (+257 * 256: the second of two tile-local levels has at most 256
    // segments, each with a partial tile)
    c.mat = 257 * (n / tile + 1) + sum_nb + c.seg + 256 + 257 * 256;
    c.part = c.mat / kScanChunk + 2;
    c.leaves = std::max<uint64_t>(1, nseg_in);
    // large leaves per level (or per segmented batch): each above one
    // 512-thread work-group's capacity
    c.huge = n / ((uint64_t)leaf_cap(dtype) + 1) + 2;
    // fill tasks per level: ceil(size/chunk) per equality bucket
    c.fill = n / kFillChunk + c.bkt + 1 + kSplStride;  // (+: two levels' equality buckets filled at once)
    return c;
}

// Parameters a sort of n keys of this type runs with: the tuned (or set)
// parameters, the leaf size capped by what one work-group holds in LDS.
gqs_params Ctx::run_params(int dtype, uint64_t n) const {
    gqs_params p = effective_params(dtype, n);
    p.leaf_max = std::min(p.leaf_max, 2 * leaf_cap(dtype));
    return p;
}

// Capacities for any sort of at most n keys (and nseg_in arrays) of this
// type, whichever size class's parameters it runs with: capacities grow with
// n, so the maximum over the parameter sets at n bounds every smaller sort.
Caps Ctx::caps_upto(uint64_t n, uint32_t nseg_in, int dtype) const {
    auto widen = [](Caps& a, const Caps& b) {
        a.seg = std::max(a.seg, b.seg);
        a.mat = std::max(a.mat, b.mat);
        a.bkt = std::max(a.bkt, b.bkt);
        a.part = std::max(a.part, b.part);
        a.leaves = std::max(a.leaves, b.leaves);
        a.fill = std::max(a.fill, b.fill);
        a.huge = std::max(a.huge, b.huge);
    };
    const uint32_t cap = 2 * leaf_cap(dtype);
    gqs_params p = params;
    p.leaf_max = std::min(p.leaf_max, cap);
    Caps c = caps_for(n, nseg_in, p, dtype);
    widen(c, caps_for(n, nseg_in, run_params(dtype, n), dtype));
    if (use_tuned)
        for (const auto& kv : tuned)
            if (kv.first.first == dtype) {
                gqs_params q = kv.second;
                q.leaf_max = std::min(q.leaf_max, cap);
                widen(c, caps_for(n, nseg_in, q, dtype));
            }
    return c;
}

gqs_status Ctx::ensure_workspace(Workspace& ws, uint64_t n, uint32_t nseg_in, const gqs_params& p, int dtype) {
    return ensure_caps(ws, caps_for(n, nseg_in, p, dtype), dtype);
}

gqs_status Ctx::ensure_caps(Workspace& ws, const Caps& need, int dtype) {
    const size_t kb = key_bytes(dtype);
    if (ws.ctl && need.seg <= ws.cap.seg && need.mat <= ws.cap.mat && need.bkt <= ws.cap.bkt &&
        need.part <= ws.cap.part && need.leaves <= ws.cap.leaves && need.fill <= ws.cap.fill &&
        need.huge <= ws.cap.huge && kb <= ws.key_bytes)
        return GQS_OK;
    Caps c = need;
    if (ws.ctl) {  // grow: keep every capacity the workspace already had
        c.seg = std::max(c.seg, ws.cap.seg);
        c.mat = std::max(c.mat, ws.cap.mat);
        c.bkt = std::max(c.bkt, ws.cap.bkt);
        c.part = std::max(c.part, ws.cap.part);
        c.leaves = std::max(c.leaves, ws.cap.leaves);
        c.fill = std::max(c.fill, ws.cap.fill);
        c.huge = std::max(c.huge, ws.cap.huge);
        // earlier calls on any stream may still use the buffers
        hipDeviceSynchronize();
        ws.release();
    }
    // Grow with headroom so that nearby sizes reuse the workspace.
    c.seg += c.seg / 4;
    c.mat += c.mat / 4;
    c.bkt += c.bkt / 4;
    c.part += c.part / 4 + 1;
    c.fill += c.fill / 4;
    hipError_t e = hipSuccess;
    if (!e) e = dalloc(&ws.segs[0], c.seg);
    if (!e) e = dalloc(&ws.segs[1], c.seg);
    if (!e) e = dalloc(&ws.gather, c.seg);
    if (!e) e = dalloc(&ws.plans, c.seg);
    if (!e) e = hipMalloc(&ws.spl, std::max<size_t>(1, c.seg * kSplStride * 8));
    if (!e) e = dalloc(&ws.meta, c.seg * kSplStride);
    if (!e) e = dalloc(&ws.lut, c.seg * kLutCells);
    if (!e) e = dalloc(&ws.mat, c.mat * kMatRegions);
    if (!e) e = dalloc(&ws.bsum, c.seg * kSplStride);
    if (!e) e = dalloc(&ws.part, c.part);
    if (!e) e = dalloc(&ws.bkt, c.bkt);
    if (!e) e = dalloc(&ws.fill, c.fill);
    if (!e) e = dalloc(&ws.ctl, (size_t)kCtlBlocks);
    if (!e) e = dalloc(&ws.leaves, c.leaves);
    if (!e) e = dalloc(&ws.huge, c.huge);
    if (e) {
        ws.release();
        fail(e, "workspace allocation");
        return GQS_ENOMEM;
    }
    ws.cap = c;
    ws.key_bytes = 8;
    return GQS_OK;
}

// Tuned parameters of the size class floor(log2 n) of this key type, or of
// the nearest tuned class of the same key type (the lower one on a tie).
gqs_params Ctx::effective_params(int dtype, uint64_t n) const {
    if (use_tuned && n > 1) {
        int lg = 0;
        while ((2ull << lg) <= n) ++lg;
        auto hi = tuned.lower_bound(std::make_pair(dtype, lg));
        auto best = tuned.end();
        if (hi != tuned.end() && hi->first.first == dtype) best = hi;
        if (hi != tuned.begin()) {
            auto lo = std::prev(hi);
            if (lo->first.first == dtype &&
                (best == tuned.end() || lg - lo->first.second <= best->first.second - lg))
                best = lo;
        }
        if (best != tuned.end()) return best->second;
    }
    return params;
}

Grid Ctx::grid(const gqs_params& p) const {
    Grid g;
    g.part = cus * (int)p.part_wgs_per_cu;
    g.splitters = std::min(1024, cus * 4);
    g.scan = cus * 2;
    g.plan = cus * 2;
    g.leaf = cus * (int)p.leaf_wgs_per_cu;
    g.leaf_big = cus * 2;  // when it claims leaves from a counter: no more work-groups than fit
    g.huge = cus;
    g.tilepart = cus;
    return g;
}

void Ctx::prof_begin(int kind, hipStream_t st) {
    if (!profiling || capturing_now) return;
    ProfRec r;
    r.kind = kind;
    hipEventCreate(&r.a);
    hipEventCreate(&r.b);
    hipEventRecord(r.a, st);
    prof.push_back(r);
}
void Ctx::prof_end(hipStream_t st) {
    if (!profiling || capturing_now || prof.empty()) return;
    hipEventRecord(prof.back().b, st);
}

uint32_t rescue_leaf_target(const gqs_params& p) { return std::max<uint32_t>(1, std::min(p.leaf_target, p.leaf_max / 2)); }

// Levels of a sort of n keys.  Multipivot: levels_needed (the expected depth
// plus a spare).  Classic (one median-of-3 pivot per segment, the reference's
// schedule): the expected classic depth (capped by max_levels), followed by
// multipivot rescue levels sized for a segment of all n keys, so that every
// bucket reaches the bounded local sort whatever depth the median-of-3 pivots
// need on the input (the reference re-launches until its work list is empty,
// CL20/QuicksortKernels.cl:132-222); all are enqueued up front and empty
// levels exit at once.
LevelPlan plan_levels(uint64_t n, const gqs_params& p) {
    LevelPlan lp{levels_needed(n, p), 0, p.leaf_target};
    if (p.mode == MODE_CLASSIC && lp.L > 0) {
        gqs_params q = p;
        q.mode = MODE_MULTIPIVOT;
        q.leaf_target = rescue_leaf_target(p);
        q.max_levels = kMaxLevels - 1;
        const uint32_t rescue = levels_needed(n, q);
        // without the classic spare; the rescue levels always fit
        lp.classic = std::max<uint32_t>(1, std::min<uint32_t>(lp.L - 1, kMaxLevels - 1 - rescue));
        lp.L = std::min<uint32_t>(lp.classic + rescue, kMaxLevels - 1);
        lp.rescue_leaf_target = q.leaf_target;
    }
    return lp;
}

uint32_t levels_needed(uint64_t n, const gqs_params& p) {
    if (n <= p.leaf_max) return 0;
    uint32_t L = 0;
    double sz = (double)n;
    const double shrink = (p.mode == MODE_CLASSIC) ? 1.5 : std::max(2.0, p.max_fanout / 2.0);
    while (sz > p.leaf_max && L < (uint32_t)kMaxLevels - 2) {
        sz = sz / shrink;
        ++L;
    }
    return std::max<uint32_t>(1, std::min<uint32_t>(L + 1, std::min<uint32_t>(p.max_levels, kMaxLevels - 1)));
}

// The whole sort as one stream-ordered chain.  Either a single key range
// [0, n) of d (offsets == nullptr) or nseg arrays delimited by offsets.
gqs_status Ctx::run(Workspace& ws, int dtype, void* d, void* dn, uint64_t n, const uint64_t* offsets,
                    uint32_t nseg_in, const uint32_t* nseg_dev, hipStream_t st, bool capturing, bool keep_status) {
    gqs_params p = run_params(dtype, n);
    if (capturing) p.host_driven = 0;
    if (gqs_status s = ensure_workspace(ws, n, nseg_in, p, dtype)) return s;
    capturing_now = capturing;
    const bool prof_on = profiling && !capturing;
    if (prof_on) {
        for (auto& r : prof) {
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
        prof.clear();
    }
    const Grid g = grid(p);
    uint32_t L;
    if (offsets) {
        L = plan_levels(n, p).L;
        prof_begin(7, st);
        if (fail(launch_init(ws.ctl, kCtlBlocks, keep_status, ws.segs[0], ws.leaves, 0, 0, 0, st), "init"))
            return GQS_EDEVICE;
        const SeedCaps sc{(uint32_t)ws.cap.seg, (uint32_t)ws.cap.leaves, (uint32_t)ws.cap.huge};
        if (fail(launch_seed_segments(offsets, nseg_in, nseg_dev, n, ws.segs[0], &ws.ctl[0], &ws.ctl[kSeedCtl],
                                      ws.leaves, p.leaf_max, leaf_cap(dtype), ws.huge, sc, st),
                 "seed"))
            return GQS_EDEVICE;
        prof_end(st);
    } else {
        const bool as_leaf = n <= p.leaf_max;
        const bool as_huge = as_leaf && n > leaf_cap(dtype);
        L = as_leaf ? 0 : plan_levels(n, p).L;
        prof_begin(7, st);
        if (fail(launch_init(ws.ctl, kCtlBlocks, keep_status, ws.segs[0], as_huge ? ws.huge : ws.leaves, 1, n,
                             SEG_RAW | (as_huge ? kInitHuge : (as_leaf ? kInitLeaf : 0u)), st),
                 "init"))
            return GQS_EDEVICE;
        prof_end(st);
    }
    LevelArgs a{};
    a.d = d;
    a.dn = dn;
    a.plans = ws.plans;
    a.spl = ws.spl;
    a.meta = ws.meta;
    a.lut = ws.lut;
    a.mat = ws.mat;
    a.bsum = ws.bsum;
    a.part = ws.part;
    a.bkt = ws.bkt;
    a.fill = ws.fill;
    a.fill_cap = (uint32_t)ws.cap.fill;
    a.mode = (uint32_t)p.mode;
    a.leaf_max = p.leaf_max;
    a.leaf_target = p.leaf_target;
    a.oversample = p.oversample;
    a.max_fanout = p.max_fanout;
    a.seg_cap = (uint32_t)ws.cap.seg;
    if (queue_limit && !capturing) a.seg_cap = std::min(a.seg_cap, queue_limit);
    a.status = &ws.ctl[kSeedCtl].overflow;
    a.mat_cap = (uint32_t)ws.cap.mat;
    a.bkt_cap = (uint32_t)ws.cap.bkt;
    a.seed = 0x5EEDu;
    a.leaf_cap = leaf_cap(dtype);
    a.huge = ws.huge;
    a.huge_cap = (uint32_t)ws.cap.huge;
    a.key_bytes = (uint32_t)key_bytes(dtype);
    // Level-0 leaves: the whole input when it fits one local sort, or the
    // short arrays of a segmented sort (all still in caller representation).
    if (offsets || L == 0) {
        LeafArgs la{};
        la.d = d;
        la.dn = dn;
        la.list = ws.leaves;
        la.list_count = &ws.ctl[kSeedCtl].nleaf_list;
        la.list_cap = (uint32_t)ws.cap.leaves;
        la.lv = a;
        la.lv.ctl = &ws.ctl[kSeedCtl];
        la.leaf_max = p.leaf_max;
        la.leaf_cap = leaf_cap(dtype);
        la.huge = ws.huge;
        la.huge_count = &ws.ctl[kSeedCtl].nhuge;
        la.is_last = 1;
        la.sort_threshold = p.sort_threshold;
        la.bitonic_cutoff = p.bitonic_cutoff;
        prof_begin(6, st);
        if (fail(launch_leaf(dtype, la, g, st), "leaf(list)")) return GQS_EDEVICE;
        prof_end(st);
    }
    const LevelPlan lplan = plan_levels(n, p);
    stats = gqs_stats{};
    stats.levels_launched = L;
    // A single range whose sort plans two partition levels and a spare runs
    // both as tile-local levels: level 0 partitions each tile in place in d
    // and hands every bucket to level 1 as a gathered segment (its runs in
    // level 0's tiles); level 1 reads those runs out of d and writes its tiles
    // to dn; the leaf pass gathers level 1's buckets from dn into d.  No count
    // pass and no scan of per-tile counts: one read and one write of the keys
    // per level.
    const bool two_local = !offsets && lplan.classic == 0 && p.mode == MODE_MULTIPIVOT && L == 3 &&
                           ws.cap.seg > (uint64_t)kSplStride && level_scheme != 1;
    uint32_t* const mat_rows = ws.mat;
    uint32_t* const mat_runs = ws.mat + ws.cap.mat;
    uint32_t* const mat_next = ws.mat + 2 * ws.cap.mat;
    for (uint32_t lvl = 0; lvl < L; ++lvl) {
        a.segs = ws.segs[lvl & 1];
        a.next_segs = ws.segs[(lvl + 1) & 1];
        a.ctl = &ws.ctl[lvl];
        a.ctl_next = &ws.ctl[lvl + 1];
        a.is_last = (lvl + 1 == L) ? 1u : 0u;
        a.seed = 0x5EEDu + 977u * lvl;
        a.mode = (lvl < lplan.classic) ? (uint32_t)MODE_CLASSIC : (uint32_t)MODE_MULTIPIVOT;
        a.leaf_target = (lvl < lplan.classic) ? p.leaf_target : lplan.rescue_leaf_target;
        a.mat = mat_rows;
        a.runlen = nullptr;
        a.src_mat = nullptr;
        a.src_pre = nullptr;
        a.gather = nullptr;
        a.next_gather = nullptr;
        // The level planned to leave only leaves (the one before the spare)
        // partitions tile-locally; the spare level, contiguous, takes any
        // bucket it left above leaf_max.
        a.tile_local = (a.mode == MODE_MULTIPIVOT && lvl + 2 == L) ? 1u : 0u;
        const int* steps = a.tile_local ? kLocalSteps : nullptr;
        int nsteps = a.tile_local ? 4 : kLevelSteps;
        const bool in_place = two_local && lvl == 0;
        if (in_place) {
            a.tile_local = 2;
            a.runlen = mat_runs;
            a.next_gather = ws.gather;
            steps = kInPlaceSteps;
            nsteps = 6;
        } else if (two_local && lvl == 1) {
            a.mat = mat_next;
            a.src_mat = mat_rows;
            a.src_pre = mat_runs;
            a.gather = ws.gather;
            a.run_tab = level_scheme == 2 ? 64u : 1024u;
            steps = kGatherSteps;
        }
        for (int k = 0; k < nsteps; ++k) {
            const int step = steps ? steps[k] : k;
            // the in-place level scans its run lengths, not its rows
            LevelArgs as = a;
            if (in_place && (step == 3 || step == 4)) as.mat = mat_runs;
            prof_begin(step_kind(step), st);
            if (fail(launch_level_step(step, dtype, as, g, st), "level")) return GQS_EDEVICE;
            prof_end(st);
        }
        if (in_place) continue;  // no leaves: every bucket is a segment of level 1
        LeafArgs la{};
        la.d = d;
        la.dn = dn;
        la.list = nullptr;
        la.list_count = nullptr;
        la.lv = a;
        la.leaf_max = p.leaf_max;
        la.leaf_cap = leaf_cap(dtype);
        la.huge = ws.huge;
        la.huge_count = &ws.ctl[lvl].nhuge;
        la.is_last = a.is_last;
        la.sort_threshold = p.sort_threshold;
        la.bitonic_cutoff = p.bitonic_cutoff;
        prof_begin(6, st);
        if (fail(launch_leaf(dtype, la, g, st), "leaf")) return GQS_EDEVICE;
        prof_end(st);
        if (p.host_driven && !a.is_last) {
            uint32_t next = 0;
            if (fail(hipMemcpyAsync(&next, &ws.ctl[lvl + 1].nseg, sizeof next, hipMemcpyDeviceToHost, st),
                     "level readback") ||
                fail(hipStreamSynchronize(st), "level sync"))
                return GQS_EDEVICE;
            stats.levels_used = lvl + 1;
            if (next == 0) break;
        }
    }
    return GQS_OK;
}

gqs_status Ctx::fetch_stats(Workspace& w, gqs_stats* out, uint32_t* bits) {
    *bits = 0;
    if (!w.ctl) return GQS_OK;
    std::vector<LevelCtl> c(kCtlBlocks);
    if ((w.done && fail(hipEventSynchronize(w.done), "sync")) ||
        fail(hipMemcpy(c.data(), w.ctl, sizeof(LevelCtl) * kCtlBlocks, hipMemcpyDeviceToHost), "stats readback"))
        return GQS_EDEVICE;
    uint32_t used = 0, over = c[kSeedCtl].oversize, ovf = 0;
    for (int i = 0; i < kMaxLevels; ++i) {
        if (c[i].nseg) used = i + 1;
        over += c[i].oversize;
        ovf |= c[i].overflow;
        if (i < 8) {
            out->segments[i] = c[i].nseg;
            out->level_keys[i] = ((uint64_t)c[i].keys_hi << 32) | c[i].keys_lo;
        }
    }
    out->levels_used = std::max(out->levels_used, used);
    out->oversize_leaves = over;
    *bits = c[kSeedCtl].overflow | ovf;
    out->overflow = *bits;
    return GQS_OK;
}

gqs_status Ctx::fetch_stats() { return fetch_stats(ws, &stats, &status_bits); }

gqs_status Ctx::status_of(uint32_t bits, const char* what) {
    char buf[256];
    if (bits & (kBadOffsets | kBadCount)) {
        std::snprintf(buf, sizeof buf, "%s: %s", what,
                      (bits & kBadOffsets) ? "an array's offsets decrease or pass total_keys (array skipped)"
                                           : "more arrays than the graph was created for (extra arrays skipped)");
        return set_error(GQS_EINVAL, buf);
    }
    if (bits & kOvfCapacity) {
        std::snprintf(buf, sizeof buf, "%s: a queue bound was exceeded (flags 0x%x); keys left unsorted", what, bits);
        return set_error(GQS_EOVERFLOW, buf);
    }
    return GQS_OK;
}

}  // namespace gqs

// ================================================================== C ABI
extern "C" {

gqs_status gqs_default_params(int dtype, gqs_params* out) {
    if (!out) return GQS_EINVAL;
    gqs_params p{};
    p.mode = GQS_MODE_MULTIPIVOT;
    p.leaf_max = 2 * leaf_cap(dtype);
    p.leaf_target = (dtype == DT_F64) ? 3072 : 6144;  // 3/8 of leaf_max: buckets near 6 Ki (f64 3 Ki) keys
    p.oversample = 16;
    p.max_fanout = 256;
    p.max_levels = 8;
    p.host_driven = 0;
    p.part_wgs_per_cu = 4;
    p.leaf_wgs_per_cu = 4;
    p.sort_threshold = (dtype == DT_F64) ? 8 : 16;
    p.bitonic_cutoff = 512;
    *out = p;
    return GQS_OK;
}

gqs_status gqs_create(gqs_ctx** out, int device) {
    if (!out) return GQS_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return GQS_EDEVICE;
    auto* c = new (std::nothrow) gqs_ctx();
    if (!c) return GQS_ENOMEM;
    c->device = device;
    hipDeviceProp_t prop{};
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return GQS_EDEVICE;
    }
    // the allocation-free entry point's scratch, looked up once per device
    // with the device current (a failure only disables that entry point)
    gqs::noalloc_setup(device, &c->noalloc_scratch);
    c->cus = prop.multiProcessorCount;
    c->lds_per_wg = (int)prop.sharedMemPerBlock;
    c->wave = prop.warpSize;
    c->name = prop.gcnArchName;
    gqs_default_params(DT_U32, &c->params);
    *out = c;
    return GQS_OK;
}

gqs_status gqs_destroy(gqs_ctx* c) {
    if (!c) return GQS_EINVAL;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        hipSetDevice(c->device);
        hipDeviceSynchronize();
        for (auto& r : c->prof) {
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
        c->ws.release_all();
        c->release_lanes();
        hipFree(c->host_d);
        hipFree(c->host_dn);
        hipFree(c->merge_split);
    }
    delete c;
    return GQS_OK;
}

const char* gqs_last_error(const gqs_ctx* c) { return c ? c->err.c_str() : "null context"; }

gqs_status gqs_get_params(const gqs_ctx* c, gqs_params* out) {
    if (!c || !out) return GQS_EINVAL;
    *out = c->params;
    return GQS_OK;
}

gqs_status gqs_check_params(const gqs_params* p) { return (p && gqs::params_valid(*p)) ? GQS_OK : GQS_EINVAL; }

gqs_status gqs_set_params(gqs_ctx* c, const gqs_params* p) {
    if (!c || !p) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!gqs::params_valid(*p)) return c->set_error(GQS_EINVAL, "gqs_set_params: parameter out of range");
    c->params = *p;
    c->use_tuned = false;
    return GQS_OK;
}

}  // extern "C"

namespace gqs {

gqs_status Ctx::sort_device_locked(int dtype, void* d, void* dn, uint64_t n, hipStream_t st) {
    if (fail(hipSetDevice(device), "hipSetDevice")) return GQS_EDEVICE;
    if (gqs_status s = stream_enter(ws, st)) return s;
    gqs_status s;
    if (n >= kMaxRunKeys || (large_run && n > large_run))
        s = sort_large(dtype, d, dn, n, large_run ? large_run : kLargeRunKeys, st);
    else
        s = run(ws, dtype, d, dn, n, nullptr, 0, nullptr, st, false);
    if (gqs_status t = stream_leave(ws, st)) return s ? s : t;
    return s;
}

bool params_valid(const gqs_params& p) {
    const bool pow2 = p.max_fanout >= 2 && p.max_fanout <= 256 && (p.max_fanout & (p.max_fanout - 1)) == 0;
    return (p.mode == GQS_MODE_MULTIPIVOT || p.mode == GQS_MODE_CLASSIC) && p.leaf_max >= 16 &&
           p.leaf_max <= 2 * leaf_cap(DT_U32) && p.leaf_target >= 1 && p.oversample >= 1 && pow2 &&
           p.max_levels >= 1 && p.max_levels <= (uint32_t)kMaxLevels - 1 && p.host_driven <= 1 &&
           p.part_wgs_per_cu >= 1 && p.part_wgs_per_cu <= 32 && p.leaf_wgs_per_cu >= 1 && p.leaf_wgs_per_cu <= 32 &&
           p.sort_threshold >= 4 && p.sort_threshold <= 4096 && p.bitonic_cutoff >= 64 && p.bitonic_cutoff <= 512;
}

}  // namespace gqs

extern "C" {

gqs_status gqs_sort_device(gqs_ctx* c, int dtype, void* d, void* dn, size_t n, void* stream) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (dtype < DT_U32 || dtype > DT_F64) return c->set_error(GQS_EINVAL, "gqs_sort_device: bad dtype");
    if (n < 2) return GQS_OK;
    if (!d || !dn) return c->set_error(GQS_EINVAL, "gqs_sort_device: null buffer");
    return c->sort_device_locked(dtype, d, dn, n, (hipStream_t)stream);
}

gqs_status gqs_set_queue_limit(gqs_ctx* c, uint32_t segments) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    c->queue_limit = segments;
    return GQS_OK;
}

gqs_status gqs_set_level_scheme(gqs_ctx* c, int scheme) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (scheme < 0 || scheme > 2) return c->set_error(GQS_EINVAL, "gqs_set_level_scheme: scheme must be 0, 1 or 2");
    c->level_scheme = (uint32_t)scheme;
    return GQS_OK;
}

gqs_status gqs_set_large_run(gqs_ctx* c, size_t keys) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (keys != 0 && (keys < 2 || keys >= gqs::kMaxRunKeys))
        return c->set_error(GQS_EINVAL, "gqs_set_large_run: run length out of range");
    c->large_run = keys;
    return GQS_OK;
}

namespace gqs {

// Runs of `run_keys` keys are sorted in place (dn's matching range is their
// scratch); then rounds of pairwise merges double the run length, each round
// reading one of d/dn and writing the other, with a lone last run copied
// across; a final copy brings the keys back to d after an odd number of rounds.
gqs_status Ctx::sort_large(int dtype, void* d, void* dn, uint64_t n, uint64_t run_keys, hipStream_t st) {
    const size_t kb = key_bytes(dtype);
    char* const d0 = static_cast<char*>(d);
    char* const dn0 = static_cast<char*>(dn);
    for (uint64_t off = 0; off < n; off += run_keys) {
        const uint64_t m = std::min(run_keys, n - off);
        if (m < 2) continue;
        if (gqs_status s = run(ws, dtype, d0 + off * kb, dn0 + off * kb, m, nullptr, 0, nullptr, st, false, off > 0))
            return s;
    }
    if (gqs_status s = ensure_merge_split(dtype, n)) return s;
    char* src = d0;
    char* dst = dn0;
    for (uint64_t w = run_keys; w < n; w *= 2) {
        for (uint64_t a = 0; a < n; a += 2 * w) {
            const uint64_t na = std::min(w, n - a), nb = std::min(w, n - a - na);
            if (nb == 0) {
                if (fail(hipMemcpyAsync(dst + a * kb, src + a * kb, na * kb, hipMemcpyDeviceToDevice, st), "merge copy"))
                    return GQS_EDEVICE;
            } else if (fail(launch_merge(dtype, src + a * kb, na, src + (a + na) * kb, nb, dst + a * kb, merge_split, st),
                            "merge")) {
                return GQS_EDEVICE;
            }
        }
        std::swap(src, dst);
    }
    if (src != d0 && fail(hipMemcpyAsync(d0, src, n * kb, hipMemcpyDeviceToDevice, st), "merge copy"))
        return GQS_EDEVICE;
    return GQS_OK;
}

gqs_status Ctx::ensure_merge_split(int dtype, uint64_t n) {
    const uint64_t need = merge_split_entries(dtype, n);
    if (need <= merge_split_cap) return GQS_OK;
    hipFree(merge_split);
    merge_split = nullptr;
    merge_split_cap = 0;
    if (hipMalloc(&merge_split, need * sizeof(uint64_t)) != hipSuccess)
        return set_error(GQS_ENOMEM, "merge scratch");
    merge_split_cap = need;
    return GQS_OK;
}

}  // namespace gqs

gqs_status gqs_set_host_pipeline(gqs_ctx* c, size_t min_keys) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (min_keys != 0 && min_keys < (1u << 16))
        return c->set_error(GQS_EINVAL, "gqs_set_host_pipeline: min_keys must be 0 or >= 2^16");
    c->host_pipe_min = min_keys;
    return GQS_OK;
}

gqs_status gqs_sort_segmented(gqs_ctx* c, int dtype, void* d, void* dn, const uint64_t* offsets, size_t nseg,
                              size_t total_keys, void* stream) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (dtype < DT_U32 || dtype > DT_F64) return c->set_error(GQS_EINVAL, "gqs_sort_segmented: bad dtype");
    if (nseg == 0 || total_keys == 0) return GQS_OK;
    if (!d || !dn || !offsets) return c->set_error(GQS_EINVAL, "gqs_sort_segmented: null buffer");
    if (nseg >= 0xFFFFFFFFull) return c->set_error(GQS_EINVAL, "gqs_sort_segmented: too many arrays");
    if (total_keys >= gqs::kMaxRunKeys)
        return c->set_error(GQS_EINVAL, "gqs_sort_segmented: total_keys must be < 2^32 - 2^16");
    if (c->fail(hipSetDevice(c->device), "hipSetDevice")) return GQS_EDEVICE;
    hipStream_t st = (hipStream_t)stream;
    if (gqs_status s = c->stream_enter(c->ws, st)) return s;
    gqs_status s = c->run(c->ws, dtype, d, dn, total_keys, offsets, (uint32_t)nseg, nullptr, st, false);
    if (gqs_status t = c->stream_leave(c->ws, st)) return s ? s : t;
    return s;
}

namespace gqs {

constexpr size_t kStageChunk = 16ull << 20;  // bytes per pinned staging chunk
constexpr int kHostLanes = 4;                 // copy lanes (host threads, streams)
constexpr uint64_t kPipeMinKeys = 1ull << 24;  // default smallest host sort that overlaps its copies
constexpr int kPipeLevels = 3;                 // merge levels of the overlapped host sort
constexpr int kPipeRuns = 1 << kPipeLevels;    // runs uploaded and sorted in turn
constexpr size_t kPinnedChunk = 64ull << 20;   // bytes per DMA when the host memory is pinned
constexpr int kPipePieces = 8;                 // output ranges of the final merge, downloaded in turn

void Ctx::release_lanes() {
    for (HostLane& l : lanes) {
        if (l.st) hipStreamSynchronize(l.st);
        for (int b = 0; b < 2; ++b) {
            if (l.ev[b]) hipEventDestroy(l.ev[b]);
            if (l.buf[b]) hipHostFree(l.buf[b]);
        }
        if (l.st) hipStreamDestroy(l.st);
    }
    lanes.clear();
    if (pipe_st) {
        hipStreamSynchronize(pipe_st);
        hipStreamDestroy(pipe_st);
        pipe_st = nullptr;
    }
    for (hipEvent_t ev : pipe_ev) hipEventDestroy(ev);
    pipe_ev.clear();
    for (hipEvent_t ev : chunk_ev) hipEventDestroy(ev);
    chunk_ev.clear();
}

// True if [p, p + bytes) is pinned host memory (the DMA engines reach it
// directly).  A pageable pointer makes the query fail; its error is cleared.
static bool host_range_pinned(const char* p, size_t bytes) {
    for (const char* q : {p, p + bytes - 1}) {
        hipPointerAttribute_t a{};
        if (hipPointerGetAttributes(&a, q) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        if (a.type != hipMemoryTypeHost) return false;
    }
    return true;
}

static size_t copy_chunk(const char* host, size_t bytes) {
    return host_range_pinned(host, bytes) ? kPinnedChunk : kStageChunk;
}

// Copies between host memory and the device in copy_chunk() pieces.  Chunks
// go round-robin to kHostLanes lanes (one lane for pinned memory).  Pageable
// memory goes through each lane's two pinned staging buffers: a lane copies
// one chunk on the host while the DMA engine moves its other one
// (host->device: fill, then copy; device->host: the next chunk's DMA is
// issued before the current chunk is drained).  Pinned memory is copied by
// the DMA engines directly.  Returns when every byte has arrived.
gqs_status Ctx::host_copy(char* dev, char* host, size_t bytes, bool to_device, const CopyHooks* hooks) {
    if (lanes.empty()) {
        lanes.resize(kHostLanes);
        for (HostLane& l : lanes) {
            hipError_t e = hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking);
            for (int b = 0; b < 2 && e == hipSuccess; ++b) {
                e = hipEventCreateWithFlags(&l.ev[b], hipEventDisableTiming);
                if (e == hipSuccess) e = hipHostMalloc(&l.buf[b], kStageChunk, hipHostMallocDefault);
            }
            if (e != hipSuccess) {
                release_lanes();
                return set_error(GQS_ENOMEM, "host staging buffers");
            }
        }
    }
    const bool pinned = bytes && host_range_pinned(host, bytes);
    const size_t csz = pinned ? kPinnedChunk : kStageChunk;
    const size_t nch = (bytes + csz - 1) / csz;
    // pinned: one lane, so chunks arrive in order at the link's full rate
    const int nl = pinned ? 1 : (int)std::min<size_t>(kHostLanes, nch);
    std::atomic<int> bad{0};
    auto lane = [&](int li) {
        if (hipSetDevice(device) != hipSuccess) {
            bad = 1;
            return;
        }
        HostLane& L = lanes[li];
        auto chunk = [&](size_t k, size_t* off, size_t* len) {
            *off = ((size_t)li + k * (size_t)nl) * csz;
            *len = std::min(csz, bytes - *off);
        };
        const size_t cnt = (nch - (size_t)li + nl - 1) / nl;  // chunks of this lane
        hipError_t e = hipSuccess;
        size_t off, len;
        if (pinned) {
            for (size_t k = 0; k < cnt && e == hipSuccess; ++k) {
                chunk(k, &off, &len);
                if (hooks && hooks->before) e = hooks->before(L.st, off, len);
                if (e == hipSuccess)
                    e = to_device ? hipMemcpyAsync(dev + off, host + off, len, hipMemcpyHostToDevice, L.st)
                                  : hipMemcpyAsync(host + off, dev + off, len, hipMemcpyDeviceToHost, L.st);
                if (e == hipSuccess && hooks && hooks->after) e = hooks->after(L.st, off, len);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(L.st);
        } else if (to_device) {
            for (size_t k = 0; k < cnt && e == hipSuccess; ++k) {
                const int b = (int)(k & 1);
                chunk(k, &off, &len);
                if (k >= 2) e = hipEventSynchronize(L.ev[b]);  // the DMA out of this buffer is done
                if (e != hipSuccess) break;
                std::memcpy(L.buf[b], host + off, len);
                if (hooks && hooks->before) e = hooks->before(L.st, off, len);
                if (e == hipSuccess) e = hipMemcpyAsync(dev + off, L.buf[b], len, hipMemcpyHostToDevice, L.st);
                if (e == hipSuccess) e = hipEventRecord(L.ev[b], L.st);
                if (e == hipSuccess && hooks && hooks->after) e = hooks->after(L.st, off, len);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(L.st);
        } else {
            auto issue = [&](size_t k) {
                size_t o, n;
                chunk(k, &o, &n);
                hipError_t r = (hooks && hooks->before) ? hooks->before(L.st, o, n) : hipSuccess;
                if (r == hipSuccess) r = hipMemcpyAsync(L.buf[k & 1], dev + o, n, hipMemcpyDeviceToHost, L.st);
                if (r == hipSuccess) r = hipEventRecord(L.ev[k & 1], L.st);
                if (r == hipSuccess && hooks && hooks->after) r = hooks->after(L.st, o, n);
                return r;
            };
            if (cnt) e = issue(0);
            for (size_t k = 0; k < cnt && e == hipSuccess; ++k) {
                if (k + 1 < cnt) e = issue(k + 1);  // its buffer was drained at step k - 1
                if (e == hipSuccess) e = hipEventSynchronize(L.ev[k & 1]);
                if (e != hipSuccess) break;
                chunk(k, &off, &len);
                std::memcpy(host + off, L.buf[k & 1], len);
            }
        }
        if (e != hipSuccess) bad = 1;
    };
    std::vector<std::thread> pool;
    for (int li = 1; li < nl; ++li) pool.emplace_back(lane, li);
    lane(0);
    for (auto& t : pool) t.join();
    return bad ? set_error(GQS_EDEVICE, to_device ? "host->device staging copy" : "device->host staging copy")
               : GQS_OK;
}

// Host sort whose copies overlap the sort (2^16 <= n < kMaxRunKeys; the
// caller holds the lock and host_d/host_dn hold n keys each).  A sort's first
// output key depends on its last input key, so the upload and the download
// cannot overlap each other; what overlaps them is the sorting.  The upload
// is one continuous chunked copy on a host thread; each chunk's DMA records
// its own event.  As soon as every chunk of run r has been issued, run r's
// sort is enqueued on pipe_st behind those chunks' events, so the GPU sorts
// it while later chunks are in flight.  The runs are merged in a binary tree
// as soon as both halves of a block are sorted (level j reads d for odd j,
// dn for even j, and writes the other), so after the last upload only the
// last run's sort and one merge per level remain.  The final merge is
// enqueued as kPipePieces ranges of whole merge tiles, each followed by an
// event, and the download (one continuous chunked copy) issues each chunk's
// DMA behind the events of the ranges it covers, so it proceeds while later
// ranges are merged.
gqs_status Ctx::sort_host_pipelined(int dtype, char* host, uint64_t n) {
    const size_t kb = key_bytes(dtype);
    const size_t bytes = n * kb;
    char* const buf[2] = {static_cast<char*>(host_d), static_cast<char*>(host_dn)};
    if (!pipe_st && fail(hipStreamCreateWithFlags(&pipe_st, hipStreamNonBlocking), "pipeline stream"))
        return GQS_EDEVICE;
    const size_t csz = copy_chunk(host, bytes);  // host_copy's chunks, for this host array
    const size_t nch = (bytes + csz - 1) / csz;
    for (std::vector<hipEvent_t>* pool : {&pipe_ev, &chunk_ev}) {
        const size_t want = pool == &pipe_ev ? (size_t)kPipePieces : nch;
        while (pool->size() < want) {
            hipEvent_t ev = nullptr;
            if (fail(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "pipeline event")) return GQS_EDEVICE;
            pool->push_back(ev);
        }
    }
    // runs [r*w, min((r+1)*w, n)); for n >= 2^16 the last one holds >= n/8 - 7 keys
    const uint64_t w = (n + kPipeRuns - 1) / kPipeRuns;
    auto edge = [&](uint64_t r) { return std::min<uint64_t>(n, r * w); };
    if (gqs_status s = ensure_workspace(ws, w, 0, run_params(dtype, w), dtype)) return s;
    if (gqs_status s = ensure_merge_split(dtype, n)) return s;
    const uint64_t T = merge_tile_keys(dtype);
    const uint64_t tiles = (n + T - 1) / T;
    auto piece = [&](int p) { return std::min<uint64_t>(n, tiles * (uint64_t)p / kPipePieces * T); };

    // upload: chunks flagged as their DMA is issued
    struct Progress {
        std::mutex m;
        std::condition_variable cv;
        std::vector<char> issued;
        bool done = false;
    } pg;
    pg.issued.assign(nch, 0);
    CopyHooks up;
    up.after = [&](hipStream_t s, size_t off, size_t) {
        const hipError_t e = hipEventRecord(chunk_ev[off / csz], s);
        if (e == hipSuccess) {
            std::lock_guard<std::mutex> g(pg.m);
            pg.issued[off / csz] = 1;
        }
        pg.cv.notify_all();
        return e;
    };
    gqs_status up_status = GQS_OK;
    std::thread uploader([&] {
        up_status = host_copy(buf[0], host, bytes, true, &up);
        {
            std::lock_guard<std::mutex> g(pg.m);
            pg.done = true;
        }
        pg.cv.notify_all();
    });
    // waits until chunks [c0, c1) are issued; false if the upload ended first
    auto issued = [&](size_t c0, size_t c1) {
        std::unique_lock<std::mutex> g(pg.m);
        auto all = [&] {
            for (size_t c = c0; c < c1; ++c)
                if (!pg.issued[c]) return false;
            return true;
        };
        pg.cv.wait(g, [&] { return pg.done || all(); });
        return all();
    };
    // merge of blocks [lo, mid) and [mid, hi) of runs at tree level j
    auto merge = [&](int j, uint64_t lo, uint64_t mid, uint64_t hi) {
        const char* src = buf[(j - 1) & 1];
        return launch_merge(dtype, src + lo * kb, mid - lo, src + mid * kb, hi - mid, buf[j & 1] + lo * kb,
                            merge_split, pipe_st);
    };
    hipStream_t st = pipe_st;
    gqs_status s = stream_enter(ws, st);
    for (int r = 0; r < kPipeRuns && s == GQS_OK; ++r) {
        const uint64_t off = edge(r), m = edge(r + 1) - off;
        const size_t c0 = off * kb / csz, c1 = ((off + m) * kb + csz - 1) / csz;
        if (!issued(c0, c1)) {
            s = GQS_EDEVICE;  // the upload failed; its status is reported below
            break;
        }
        for (size_t c = c0; c < c1 && s == GQS_OK; ++c)
            if (fail(hipStreamWaitEvent(st, chunk_ev[c], 0), "pipeline wait")) s = GQS_EDEVICE;
        if (s == GQS_OK) s = run(ws, dtype, buf[0] + off * kb, buf[1] + off * kb, m, nullptr, 0, nullptr, st, false, r > 0);
        // every block of 2^j runs that run r completes, below the final level
        for (int j = 1; j < kPipeLevels && s == GQS_OK && ((r + 1) & ((1 << j) - 1)) == 0; ++j) {
            const uint64_t r0 = (uint64_t)r + 1 - (1ull << j);
            if (fail(merge(j, edge(r0), edge(r0 + (1ull << (j - 1))), edge(r + 1)), "merge")) s = GQS_EDEVICE;
        }
    }
    // the final level, range by range
    const uint64_t h = edge(kPipeRuns / 2);
    const char* fsrc = buf[(kPipeLevels - 1) & 1];
    char* const out = buf[kPipeLevels & 1];
    if (s == GQS_OK && fail(launch_merge_split(dtype, fsrc, h, fsrc + h * kb, n - h, merge_split, st), "merge"))
        s = GQS_EDEVICE;
    for (int p = 0; p < kPipePieces && s == GQS_OK; ++p) {
        const uint64_t t0 = tiles * (uint64_t)p / kPipePieces, t1 = tiles * (uint64_t)(p + 1) / kPipePieces;
        if (fail(launch_merge_tiles(dtype, fsrc, h, fsrc + h * kb, n - h, out, merge_split, t0, t1, st), "merge") ||
            fail(hipEventRecord(pipe_ev[p], st), "pipeline event"))
            s = GQS_EDEVICE;
    }
    const gqs_status t = stream_leave(ws, st);
    uploader.join();
    if (up_status != GQS_OK) return up_status;
    if (s != GQS_OK) return s;
    if (t != GQS_OK) return t;
    // download: each chunk's DMA waits for the merge ranges it covers
    CopyHooks down;
    down.before = [&](hipStream_t ls, size_t off, size_t len) {
        const uint64_t k0 = off / kb, k1 = (off + len) / kb;
        for (int p = 0; p < kPipePieces; ++p)
            if (piece(p) < k1 && piece(p + 1) > k0)
                if (hipError_t e = hipStreamWaitEvent(ls, pipe_ev[p], 0)) return e;
        return hipSuccess;
    };
    if (gqs_status e = host_copy(out, host, bytes, false, &down)) return e;
    // the sticky status of every run decides the result
    if (gqs_status e = fetch_stats()) return e;
    return status_of(status_bits, "gqs_sort");
}

}  // namespace gqs

gqs_status gqs_sort(gqs_ctx* c, int dtype, void* d, void* dn, size_t n) {
    if (!c) return GQS_EINVAL;
    if (dtype < DT_U32 || dtype > DT_F64) return c->set_error(GQS_EINVAL, "gqs_sort: bad dtype");
    if (n < 2) return GQS_OK;
    if (!d || !dn) return c->set_error(GQS_EINVAL, "gqs_sort: null buffer");
    const size_t bytes = n * key_bytes(dtype);
    // the whole upload / sort / download holds the lock: the device copies
    // and the staging lanes belong to the context
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->fail(hipSetDevice(c->device), "hipSetDevice")) return GQS_EDEVICE;
    if (bytes > c->host_cap) {
        hipDeviceSynchronize();
        hipFree(c->host_d);
        hipFree(c->host_dn);
        c->host_d = c->host_dn = nullptr;
        c->host_cap = 0;
        if (hipMalloc(&c->host_d, bytes) != hipSuccess || hipMalloc(&c->host_dn, bytes) != hipSuccess) {
            hipFree(c->host_d);
            c->host_d = nullptr;
            return c->set_error(GQS_ENOMEM, "gqs_sort: device buffers");
        }
        c->host_cap = bytes;
    }
    // the copy lanes' streams do not wait for earlier sorts of this context
    if (c->ws.done && c->fail(hipEventSynchronize(c->ws.done), "sync")) return GQS_EDEVICE;
    const uint64_t pipe_min = c->host_pipe_min ? c->host_pipe_min : kPipeMinKeys;
    if (n >= pipe_min && n < kMaxRunKeys && !c->large_run) return c->sort_host_pipelined(dtype, (char*)d, n);
    if (gqs_status s = c->host_copy((char*)c->host_d, (char*)d, bytes, true)) return s;
    if (gqs_status s = c->sort_device_locked(dtype, c->host_d, c->host_dn, n, nullptr)) return s;
    // synchronous: the sticky status of this call decides the result
    if (gqs_status s = c->fetch_stats()) return s;
    if (gqs_status s = c->status_of(c->status_bits, "gqs_sort")) return s;
    return c->host_copy((char*)c->host_d, (char*)d, bytes, false);
}

namespace gqs {

// One partition level of [0, n) of d into dn, without the local sort: the
// level chain of run() for a single segment, stopped after the scatter.
gqs_status Ctx::partition_level(int dtype, void* d, void* dn, uint64_t n, LevelResult* out) {
    gqs_params p = run_params(dtype, n);
    p.host_driven = 0;
    if (gqs_status s = ensure_workspace(ws, n, 0, p, dtype)) return s;
    hipStream_t st = nullptr;
    if (gqs_status s = stream_enter(ws, st)) return s;
    capturing_now = false;
    if (fail(launch_init(ws.ctl, kCtlBlocks, false, ws.segs[0], ws.leaves, 1, n, SEG_RAW, st), "init"))
        return GQS_EDEVICE;
    LevelArgs a{};
    a.d = d;
    a.dn = dn;
    a.segs = ws.segs[0];
    a.next_segs = ws.segs[1];
    a.plans = ws.plans;
    a.spl = ws.spl;
    a.meta = ws.meta;
    a.lut = ws.lut;
    a.mat = ws.mat;
    a.bsum = ws.bsum;
    a.part = ws.part;
    a.bkt = ws.bkt;
    a.fill = ws.fill;
    a.ctl = &ws.ctl[0];
    a.ctl_next = &ws.ctl[1];
    a.mode = (uint32_t)p.mode;
    a.leaf_max = p.leaf_max;
    a.leaf_target = p.leaf_target;
    a.oversample = p.oversample;
    a.max_fanout = p.max_fanout;
    a.is_last = 1;
    a.seg_cap = (uint32_t)ws.cap.seg;
    a.mat_cap = (uint32_t)ws.cap.mat;
    a.bkt_cap = (uint32_t)ws.cap.bkt;
    a.fill_cap = (uint32_t)ws.cap.fill;
    a.seed = 0x5EEDu;
    a.leaf_cap = leaf_cap(dtype);
    a.huge = ws.huge;
    a.huge_cap = (uint32_t)ws.cap.huge;
    a.status = &ws.ctl[kSeedCtl].overflow;
    a.key_bytes = (uint32_t)key_bytes(dtype);
    const Grid g = grid(p);
    for (int step = 0; step < kLevelSteps; ++step)
        if (fail(launch_level_step(step, dtype, a, g, st), "level")) return GQS_EDEVICE;
    if (gqs_status s = stream_leave(ws, st)) return s;
    const size_t kb = key_bytes(dtype);
    SegPlan plan{};
    uint16_t meta[kSplStride];
    if (fail(hipStreamSynchronize(st), "level sync") ||
        fail(hipMemcpy(&plan, ws.plans, sizeof plan, hipMemcpyDeviceToHost), "plan readback") ||
        fail(hipMemcpy(meta, ws.meta, sizeof meta, hipMemcpyDeviceToHost), "meta readback") ||
        fail(hipMemcpy(out->splitters, ws.spl, kSplStride * kb, hipMemcpyDeviceToHost), "splitter readback"))
        return GQS_EDEVICE;
    const uint32_t nb = plan.nbuckets;
    if (nb < 2 || nb > (uint32_t)kSplStride) return set_error(GQS_EDEVICE, "partition_level: bad bucket count");
    // unique splitter i maps to bucket base[i] (+1 for its equality bucket);
    // the first table entry past the splitters holds base nb - 1
    uint32_t m = 0;
    while (m < (uint32_t)kSplStride && (meta[m] & 0xFFu) != nb - 1) {
        out->eq[m] = (uint8_t)((meta[m] >> 8) & 1u);
        ++m;
    }
    out->nsplit = m;
    out->nbuckets = nb;
    std::vector<LeafRec> b(nb);
    if (fail(hipMemcpy(b.data(), ws.bkt, nb * sizeof(LeafRec), hipMemcpyDeviceToHost), "bucket readback"))
        return GQS_EDEVICE;
    for (uint32_t i = 0; i < nb; ++i) {
        out->start[i] = b[i].start;
        out->size[i] = b[i].n & ~kBucketDirect;
        out->bucket_eq[i] = (b[i].n & kBucketDirect) ? 1 : 0;
    }
    uint32_t bits = 0;
    gqs_stats tmp{};
    if (gqs_status s = fetch_stats(ws, &tmp, &bits)) return s;
    return status_of(bits, "gqs_partition_level");
}

}  // namespace gqs

gqs_status gqs_partition_level(gqs_ctx* c, int dtype, void* d, void* dn, size_t n, void* splitters, uint8_t* eq,
                               uint32_t* nsplit, uint64_t* bucket_start, uint32_t* bucket_size, uint8_t* bucket_eq,
                               uint32_t* nbuckets) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (dtype < DT_U32 || dtype > DT_F64) return c->set_error(GQS_EINVAL, "gqs_partition_level: bad dtype");
    if (!d || !dn || !splitters || !eq || !nsplit || !bucket_start || !bucket_size || !bucket_eq || !nbuckets)
        return c->set_error(GQS_EINVAL, "gqs_partition_level: null pointer");
    if (n < 2 || n >= gqs::kMaxRunKeys) return c->set_error(GQS_EINVAL, "gqs_partition_level: n out of range");
    if (c->fail(hipSetDevice(c->device), "hipSetDevice")) return GQS_EDEVICE;
    const size_t kb = key_bytes(dtype);
    std::vector<unsigned char> spl(gqs::kSplStride * kb);
    gqs::LevelResult r{};
    std::vector<uint64_t> st(gqs::kSplStride);
    std::vector<uint32_t> sz(gqs::kSplStride);
    std::vector<uint8_t> be(gqs::kSplStride);
    r.splitters = spl.data();
    r.eq = eq;
    r.start = st.data();
    r.size = sz.data();
    r.bucket_eq = be.data();
    gqs_status s = c->partition_level(dtype, d, dn, n, &r);
    if (s != GQS_OK) return s;
    std::memcpy(splitters, spl.data(), r.nsplit * kb);
    *nsplit = r.nsplit;
    *nbuckets = r.nbuckets;
    std::memcpy(bucket_start, st.data(), r.nbuckets * sizeof(uint64_t));
    std::memcpy(bucket_size, sz.data(), r.nbuckets * sizeof(uint32_t));
    std::memcpy(bucket_eq, be.data(), r.nbuckets);
    return GQS_OK;
}

gqs_status gqs_sort_each(gqs_ctx* c, int dtype, void* d, void* dn, const uint64_t* host_offsets, size_t nseg,
                         void* stream) {
    if (!c || !host_offsets) return GQS_EINVAL;
    const size_t kb = key_bytes(dtype);
    for (size_t i = 0; i < nseg; ++i) {
        const uint64_t lo = host_offsets[i], hi = host_offsets[i + 1];
        if (hi < lo) return c->set_error(GQS_EINVAL, "gqs_sort_each: offsets must be non-decreasing");
        gqs_status s = gqs_sort_device(c, dtype, static_cast<char*>(d) + lo * kb, static_cast<char*>(dn) + lo * kb,
                                       hi - lo, stream);
        if (s) return s;
    }
    return GQS_OK;
}

// ---- captured segmented sort ------------------------------------------------
struct gqs_graph {
    gqs_ctx* ctx = nullptr;
    gqs::Workspace ws;  // ws.done / ws.last order replays on different streams
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

gqs_status gqs_graph_create_segmented(gqs_ctx* c, int dtype, void* d, void* dn, const uint64_t* offsets,
                                      const uint32_t* nseg_dev, size_t max_nseg, size_t max_total_keys,
                                      gqs_graph** out) {
    if (!c || !out) return GQS_EINVAL;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    if (dtype < DT_U32 || dtype > DT_F64) return c->set_error(GQS_EINVAL, "gqs_graph_create_segmented: bad dtype");
    if (!d || !dn || !offsets || !nseg_dev || max_nseg == 0 || max_total_keys < 2 || max_nseg >= 0xFFFFFFFFull ||
        max_total_keys >= gqs::kMaxRunKeys)
        return c->set_error(GQS_EINVAL, "gqs_graph_create_segmented: bad argument");
    if (c->fail(hipSetDevice(c->device), "hipSetDevice")) return GQS_EDEVICE;
    auto* g = new (std::nothrow) gqs_graph();
    if (!g) return GQS_ENOMEM;
    g->ctx = c;
    // sized for the parameters the capture runs with (the size class of
    // max_total_keys): no allocation may happen while the stream captures
    if (gqs_status s = c->ensure_caps(g->ws, c->caps_upto(max_total_keys, (uint32_t)max_nseg, dtype), dtype)) {
        delete g;
        return s;
    }
    hipStream_t cs = nullptr;
    if (c->fail(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream")) {
        g->ws.release_all();
        delete g;
        return GQS_EDEVICE;
    }
    gqs_status st = GQS_OK;
    if (c->fail(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture")) {
        st = GQS_EDEVICE;
    } else {
        st = c->run(g->ws, dtype, d, dn, max_total_keys, offsets, (uint32_t)max_nseg, nseg_dev, cs, true);
        c->capturing_now = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &graph);
        if (st == GQS_OK && c->fail(e, "end capture")) st = GQS_EDEVICE;
        g->graph = graph;
        if (st == GQS_OK && c->fail(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0), "instantiate"))
            st = GQS_EDEVICE;
    }
    hipStreamDestroy(cs);
    if (st != GQS_OK) {
        if (g->exec) hipGraphExecDestroy(g->exec);
        if (g->graph) hipGraphDestroy(g->graph);
        g->ws.release_all();
        delete g;
        return st;
    }
    *out = g;
    return GQS_OK;
}

gqs_status gqs_graph_launch(gqs_graph* g, void* stream) {
    if (!g || !g->exec) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(g->ctx->mu);
    hipStream_t st = (hipStream_t)stream;
    if (gqs_status s = g->ctx->stream_enter(g->ws, st)) return s;
    if (g->ctx->fail(hipGraphLaunch(g->exec, st), "hipGraphLaunch")) return GQS_EDEVICE;
    return g->ctx->stream_leave(g->ws, st);
}

gqs_status gqs_graph_status(gqs_graph* g) {
    if (!g) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(g->ctx->mu);
    gqs_stats st{};
    uint32_t bits = 0;
    if (gqs_status s = g->ctx->fetch_stats(g->ws, &st, &bits)) return s;
    return g->ctx->status_of(bits, "graph replay");
}

gqs_status gqs_graph_destroy(gqs_graph* g) {
    if (!g) return GQS_EINVAL;
    hipDeviceSynchronize();
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    g->ws.release_all();
    delete g;
    return GQS_OK;
}

gqs_status gqs_get_stats(gqs_ctx* c, gqs_stats* out) {
    if (!c || !out) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (gqs_status s = c->fetch_stats()) return s;
    *out = c->stats;
    return c->status_of(c->status_bits, "last sort");
}

gqs_status gqs_set_profiling(gqs_ctx* c, int on) {
    if (!c) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    c->profiling = on != 0;
    return GQS_OK;
}

gqs_status gqs_kernel_times(gqs_ctx* c, uint32_t* launches, double* ms, int nkinds) {
    if (!c || !launches || !ms || nkinds <= 0) return GQS_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    for (int k = 0; k < nkinds; ++k) {
        launches[k] = 0;
        ms[k] = 0.0;
    }
    if (c->ws.done && c->fail(hipEventSynchronize(c->ws.done), "sync")) return GQS_EDEVICE;
    c->trace.clear();
    for (auto& r : c->prof) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
            c->trace.emplace_back(r.kind, t);
            if (r.kind < nkinds) {
                launches[r.kind]++;
                ms[r.kind] += t;
            }
        }
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    c->prof.clear();
    return GQS_OK;
}

int gqs_kernel_trace(gqs_ctx* c, int* kinds, float* ms, int cap) {
    if (!c || cap < 0 || (cap > 0 && (!kinds || !ms))) return -1;
    std::lock_guard<std::mutex> lk(c->mu);
    const int n = (int)c->trace.size();
    for (int i = 0; i < n && i < cap; ++i) {
        kinds[i] = c->trace[i].first;
        ms[i] = c->trace[i].second;
    }
    return n;
}

gqs_status gqs_device_report(gqs_ctx* c, gqs_device_props* out) {
    if (!c || !out) return GQS_EINVAL;
    hipDeviceProp_t p{};
    if (c->fail(hipGetDeviceProperties(&p, c->device), "hipGetDeviceProperties")) return GQS_EDEVICE;
    gqs_device_props r{};
    std::snprintf(r.name, sizeof r.name, "%s", p.name[0] ? p.name : p.gcnArchName);
    std::snprintf(r.arch, sizeof r.arch, "%s", p.gcnArchName);
    std::snprintf(r.pci_bus_id, sizeof r.pci_bus_id, "%04x:%02x:%02x.0", (unsigned)p.pciDomainID,
                  (unsigned)p.pciBusID, (unsigned)p.pciDeviceID);
    hipDriverGetVersion(&r.driver_version);
    hipRuntimeGetVersion(&r.runtime_version);
    r.compute_units = p.multiProcessorCount;
    r.wave_size = p.warpSize;
    r.max_threads_per_block = p.maxThreadsPerBlock;
    for (int i = 0; i < 3; ++i) r.max_block_dim[i] = p.maxThreadsDim[i];
    r.clock_khz = p.clockRate;
    r.mem_clock_khz = p.memoryClockRate;
    r.mem_bus_bits = p.memoryBusWidth;
    r.lds_per_block = p.sharedMemPerBlock;
    r.lds_per_cu = p.maxSharedMemoryPerMultiProcessor;
    r.l2_bytes = (uint64_t)p.l2CacheSize;
    r.global_mem_bytes = p.totalGlobalMem;
    *out = r;
    return GQS_OK;
}

gqs_status gqs_device_info(gqs_ctx* c, int* cus, int* lds_per_wg, int* wave_size, char* name, size_t name_len) {
    if (!c) return GQS_EINVAL;
    if (cus) *cus = c->cus;
    if (lds_per_wg) *lds_per_wg = c->lds_per_wg;
    if (wave_size) *wave_size = c->wave;
    if (name && name_len) {
        std::strncpy(name, c->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    return GQS_OK;
}

}  // extern "C"
