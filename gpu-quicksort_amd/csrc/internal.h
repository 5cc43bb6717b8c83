// internal.h -- context and workspace of the GPU-Quicksort library (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gqsort.h"
#include "kernels.h"
#include "records.h"

namespace gqs {

struct Caps {
    uint64_t seg = 0, mat = 0, bkt = 0, part = 0, leaves = 0, fill = 0, huge = 0;
};

// Largest leaf one 512-thread work-group sorts (keys); leaves up to twice
// that go to the 1024-thread instantiation through the large-leaf list.
inline uint32_t leaf_cap(int dtype) { return dtype == DT_F64 ? 4096u : 8192u; }

// Device workspace, grown on demand and reused across sorts.
struct Workspace {
    SegRec* segs[2] = {nullptr, nullptr};  // level queues (ping-pong)
    SegGather* gather = nullptr;           // runs of the gathered segments of a second tile-local level
    SegPlan* plans = nullptr;
    void* spl = nullptr;
    uint16_t* meta = nullptr;
    uint8_t* lut = nullptr;
    uint32_t* mat = nullptr;
    uint32_t* bsum = nullptr;  // keys per bucket of a tile-local level (kSplStride per segment)
    uint32_t* part = nullptr;
    LeafRec* bkt = nullptr;
    FillTask* fill = nullptr;
    LevelCtl* ctl = nullptr;
    LeafRec* leaves = nullptr;
    LeafRec* huge = nullptr;  // large-leaf list (reused by every level)
    Caps cap;
    size_t key_bytes = 0;
    // Stream order across calls: calls return once their kernels are
    // enqueued, and every call on this workspace uses the same buffers, so a
    // call on another stream first waits for `done`, recorded after the
    // previous call's last launch on `last`.
    hipStream_t last = nullptr;
    hipEvent_t done = nullptr;
    void release();      // device buffers (the stream-order event stays)
    void release_all();  // buffers and event
};

struct ProfRec {
    int kind = 0;
    hipEvent_t a = nullptr, b = nullptr;
};

// Steps of one partition level and their profiling kinds.  Step ids: 0
// splitters, 1 tileplan, 2 count, 3 scan_reduce, 4 scan_apply, 5 plan, 6
// scatter, 7 tilepart, 8 plan of a tile-local level, 9 plan of an in-place
// tile-local level, 10 tilepart of gathered segments.  A contiguous level
// runs 0-6, a tile-local level 0, 1, 7, 8 (equality buckets are filled by the
// level's leaf pass either way); a sort of two tile-local levels runs 0, 1,
// 7, 3, 4, 9 (in place, the scan over its run lengths, no leaf pass) and then
// 0, 1, 10, 8.
constexpr int kLevelSteps = 7;
constexpr int kLocalSteps[4] = {0, 1, 7, 8};
constexpr int kInPlaceSteps[6] = {0, 1, 7, 3, 4, 9};
constexpr int kGatherSteps[4] = {0, 1, 10, 8};
inline int step_kind(int step) {
    static const int kinds[11] = {0, 1, 2, 3, 3, 4, 5, 9, 4, 4, 9};
    return kinds[step];
}
// The count matrix holds three regions of Caps::mat entries: a level's own
// rows (region 0), and for a sort of two tile-local levels the first level's
// scanned run lengths (region 1) and the second level's rows (region 2),
// while the first level's rows stay in region 0.
constexpr int kMatRegions = 3;
hipError_t launch_level_step(int step, int dtype, const LevelArgs& a, const Grid& g, hipStream_t st);

// One run of the partition levels holds fewer than kMaxRunKeys keys: segment
// sizes and in-segment offsets are 32-bit, and tile bounds (an offset plus
// up to two 16 Ki-key tiles) must not wrap.  Larger sorts merge runs of
// kLargeRunKeys keys.
constexpr uint64_t kMaxRunKeys = (1ull << 32) - (1ull << 16);
constexpr uint64_t kLargeRunKeys = 1ull << 31;

Caps caps_for(uint64_t n, uint32_t nseg_in, const gqs_params& p, int dtype);

// Allocation-free sorts (gqs_sort_device_noalloc): at most kNoAllocMaxKeys
// keys, workspace carved out of a kNoAllocScratchBytes device array per GPU.
constexpr uint64_t kNoAllocMaxKeys = 1ull << 24;
constexpr size_t kNoAllocScratchBytes = 24ull << 20;
// Device address of that array on `device` (the current device), looked up
// once per device.
gqs_status noalloc_setup(int device, unsigned char** out);
// GQS_OK when every field of p is in the range the kernels assume.
bool params_valid(const gqs_params& p);
uint32_t levels_needed(uint64_t n, const gqs_params& p);
uint32_t rescue_leaf_target(const gqs_params& p);
struct LevelPlan {
    uint32_t L;                   // levels enqueued
    uint32_t classic;             // levels 0 .. classic-1 take one median-of-3 pivot (classic mode)
    uint32_t rescue_leaf_target;  // leaf target of the multipivot levels after them
};
LevelPlan plan_levels(uint64_t n, const gqs_params& p);

// Control block of the arrays a sort starts from (the single range, or the
// arrays of a segmented sort that go straight to a local sort), after the
// kMaxLevels per-level blocks; its `overflow` word is the call's sticky
// status, which every kernel that hits a bound also sets.
constexpr int kSeedCtl = kMaxLevels;
constexpr int kCtlBlocks = kMaxLevels + 1;

// LevelCtl::overflow bits.
constexpr uint32_t kOvfMatrix = 1u;     // count matrix or bucket table
constexpr uint32_t kOvfSegs = 2u;       // next level's segment queue
constexpr uint32_t kOvfFill = 4u;       // equality-bucket fill tasks
constexpr uint32_t kOvfHuge = 8u;       // large-leaf list
constexpr uint32_t kBadOffsets = 16u;   // an array's bounds are decreasing or beyond total_keys
constexpr uint32_t kBadCount = 32u;     // more arrays than the graph was created for
constexpr uint32_t kOvfCapacity = kOvfMatrix | kOvfSegs | kOvfFill | kOvfHuge;

// Result of one partition level (host memory, kSplStride entries each).
struct LevelResult {
    void* splitters;     // unique splitters, ordered-key representation
    uint8_t* eq;         // 1: the splitter has an equality bucket
    uint32_t nsplit;
    uint64_t* start;     // per bucket: first key
    uint32_t* size;      // per bucket: keys
    uint8_t* bucket_eq;  // 1: an equality bucket (its keys are not stored)
    uint32_t nbuckets;
};

struct Ctx {
    int device = 0;
    int cus = 256;
    int lds_per_wg = 65536;
    int wave = 64;
    std::string name;
    gqs_params params{};
    std::mutex mu;
    std::string err;
    Workspace ws;
    gqs_stats stats{};
    uint32_t status_bits = 0;  // sticky overflow word of the last call (after fetch_stats)
    uint32_t queue_limit = 0;  // testing hook: caps the segment queue (0: none)
    uint32_t level_scheme = 0;  // testing hook (gqs_set_level_scheme): 0 default, 1 contiguous + tile-local, 2 short run tables
    unsigned char* noalloc_scratch = nullptr;  // this device's allocation-free workspace (noalloc.hip)
    bool profiling = false;
    std::vector<ProfRec> prof;
    std::vector<std::pair<int, float>> trace;  // per launch of the last profiled sort
    // device copies for the host-pointer entry point
    void* host_d = nullptr;
    void* host_dn = nullptr;
    size_t host_cap = 0;
    // Host-pointer sorts: copy lanes, each a host thread with its own stream
    // and two pinned staging chunks (allocated on the first such sort).
    struct HostLane {
        void* buf[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        hipStream_t st = nullptr;
    };
    std::vector<HostLane> lanes;
    // Per-chunk hooks of host_copy (a chunk is kStageChunk bytes at an offset
    // that is a multiple of it): `before` runs before the chunk's DMA is
    // issued on the lane's stream, `after` right after it is issued.
    struct CopyHooks {
        std::function<hipError_t(hipStream_t, size_t off, size_t len)> before, after;
    };
    gqs_status host_copy(char* dev, char* host, size_t bytes, bool to_device, const CopyHooks* hooks = nullptr);
    void release_lanes();
    std::vector<hipEvent_t> chunk_ev;  // one per uploaded chunk of a pipelined host sort
    // Host sorts of at least host_pipe_min keys (0: the default) overlap the
    // transfers with the sort: runs sorted as they arrive, merged, and the
    // final merge's output ranges copied back as each one is written.
    uint64_t host_pipe_min = 0;
    hipStream_t pipe_st = nullptr;
    std::vector<hipEvent_t> pipe_ev;
    gqs_status ensure_merge_split(int dtype, uint64_t n);
    gqs_status sort_host_pipelined(int dtype, char* host, uint64_t n);

    bool fail(hipError_t e, const char* what);
    gqs_status set_error(gqs_status s, const char* msg);
    bool capturing_now = false;
    gqs_status ensure_workspace(Workspace& w, uint64_t n, uint32_t nseg_in, const gqs_params& p, int dtype);
    gqs_status ensure_caps(Workspace& w, const Caps& need, int dtype);
    gqs_params run_params(int dtype, uint64_t n) const;
    Caps caps_upto(uint64_t n, uint32_t nseg_in, int dtype) const;
    Grid grid(const gqs_params& p) const;
    gqs_params effective_params(int dtype, uint64_t n) const;
    std::map<std::pair<int, int>, gqs_params> tuned;  // (dtype, floor(log2 n)) -> parameters
    std::map<std::pair<int, int>, double> tuned_ms;
    bool use_tuned = false;
    void prof_begin(int kind, hipStream_t st);
    void prof_end(hipStream_t st);
    // Enqueue one sort on `st` using workspace `w`.  nseg_dev (optional) is a
    // device word holding the number of arrays, read when the chain runs.
    // keep_status: a later run of the same call (the large-n path) keeps the
    // sticky status word that earlier runs may have set.
    gqs_status run(Workspace& w, int dtype, void* d, void* dn, uint64_t n, const uint64_t* offsets,
                   uint32_t nseg_in, const uint32_t* nseg_dev, hipStream_t st, bool capturing,
                   bool keep_status = false);
    gqs_status partition_level(int dtype, void* d, void* dn, uint64_t n, LevelResult* out);
    gqs_status fetch_stats();
    gqs_status fetch_stats(Workspace& w, gqs_stats* out, uint32_t* bits);
    // Status of a call from its sticky overflow word.
    gqs_status status_of(uint32_t bits, const char* what);
    gqs_status stream_enter(Workspace& w, hipStream_t st);
    gqs_status stream_leave(Workspace& w, hipStream_t st);
    // gqs_sort_device without taking the lock (the caller holds it).
    gqs_status sort_device_locked(int dtype, void* d, void* dn, uint64_t n, hipStream_t st);
    // Sorts beyond one run's 32-bit sizes: runs of `run_keys` keys sorted by
    // run(), then merged pairwise (merge scratch grown on demand).
    uint64_t large_run = 0;  // 0: default
    uint64_t* merge_split = nullptr;
    size_t merge_split_cap = 0;
    gqs_status sort_large(int dtype, void* d, void* dn, uint64_t n, uint64_t run_keys, hipStream_t st);
};

}  // namespace gqs

struct gqs_ctx : gqs::Ctx {};
