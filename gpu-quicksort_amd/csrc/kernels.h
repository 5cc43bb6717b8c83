// kernels.h -- host-side launch interface of the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "records.h"

namespace gqs {

// Per-bucket result of one partition level (a LeafRec in the level's bucket
// table): where the bucket's keys are and how many.  Bit 31 of `n` marks an
// equality bucket: its keys all equal one splitter, the scatter does not store
// them and the level's leaf pass writes the value into d after the scatter.
constexpr uint32_t kBucketDirect = 0x80000000u;

// Everything one partition level reads and writes.  Device pointers only.
struct LevelArgs {
    void* d;                 // caller keys (in/out)
    void* dn;                // caller scratch
    const SegRec* segs;      // this level's segments
    SegRec* next_segs;       // next level's segments (appended)
    SegPlan* plans;          // per segment
    void* spl;               // splitters, kSplStride per segment (K)
    uint16_t* meta;          // bucket-id table, kSplStride per segment
    uint8_t* lut;            // splitter lookup table, kLutCells per segment
    uint32_t* mat;           // count matrix (scanned in place)
    uint32_t* bsum;          // tile-local level: keys per bucket, kSplStride per segment (k_tilepart adds)
    uint32_t* part;          // scan chunk partials
    LeafRec* bkt;            // per bucket of this level
    FillTask* fill;          // equality-bucket fill tasks of this level
    LevelCtl* ctl;           // this level
    LevelCtl* ctl_next;      // next level
    uint32_t mode;           // Mode
    uint32_t leaf_max;       // buckets up to this size are sorted locally
    uint32_t leaf_target;    // multipivot: bucket size the fan-out aims for
    uint32_t oversample;     // multipivot: sample keys per bucket
    uint32_t max_fanout;     // multipivot: largest fan-out (power of two <= 256)
    uint32_t is_last;        // 1: no next level, every bucket is sorted locally
    uint32_t seg_cap;        // capacity of next_segs
    uint32_t mat_cap;        // capacity of mat
    uint32_t bkt_cap;        // capacity of bkt
    uint32_t fill_cap;       // capacity of fill
    uint32_t seed;           // sample-position hash seed
    uint32_t leaf_cap;       // largest leaf of one 512-thread work-group (keys)
    LeafRec* huge;           // leaves above leaf_cap (and <= leaf_max), appended by k_plan
    uint32_t huge_cap;       // capacity of huge
    uint32_t tile_local;     // 1: this level partitions each tile into the same range of dn
                             // (k_tilepart); its buckets are gathered leaves, those above
                             // leaf_max rescue segments.  2: each tile in place in d, with the
                             // run lengths in runlen; every bucket becomes a gathered segment
                             // of the next level (k_plan_gather)
    uint32_t* status;        // the call's sticky overflow word (every bound hit is also ORed here)
    SegGather* next_gather;  // tile_local 2: where each next-level segment's runs are (next_segs' slots)
    const SegGather* gather; // gathered segments (SEG_GATHER): where this level's segments' runs are
    uint32_t* runlen;        // tile_local 2: keys of bucket b in tile t, laid out as mat (then scanned)
    const uint32_t* src_mat; // gathered segments: the partitioning level's in-tile starts S(b, t)
    const uint32_t* src_pre; // gathered segments: its scanned run lengths (P(b, t) by difference)
    uint32_t key_bytes;      // 4 or 8
    uint32_t run_tab;        // gathered segments: run-table entries per pass (kScatterWG; fewer in tests)
};

// Local-sort (leaf) arguments.  Either an explicit list of segments
// (list != nullptr) or every qualifying bucket of one level.
struct LeafArgs {
    void* d;
    void* dn;
    const LeafRec* list;     // explicit leaves (raw or not), count in *list_count
    const uint32_t* list_count;
    uint32_t list_cap;       // entries the list holds (a count above it was clamped and flagged)
    LevelArgs lv;            // bucket mode: the level whose buckets to sort
    uint32_t leaf_max;
    uint32_t is_last;
    uint32_t sort_threshold;   // sub-bucket size the local partition aims for
    uint32_t bitonic_cutoff;   // largest sub-bucket one wave sorts (<= 512)
    uint32_t leaf_cap;         // largest leaf of one 512-thread work-group; larger
                               // ones (<= leaf_max) come from the large-leaf list
    const LeafRec* huge;       // the large-leaf list of this pass, count in *huge_count
    const uint32_t* huge_count;
    uint32_t* next;            // the work counter of this launch (zeroed with the level's counters), or null: static order
};

struct Grid {
    int splitters;  // work-groups of the splitter kernel
    int part;       // work-groups of count / scatter
    int scan;       // work-groups of the scan kernels
    int plan;       // work-groups of the plan kernel
    int leaf;       // work-groups of the small-leaf kernel and of the leaf kernels in static order
    int leaf_big;   // work-groups of the 512-thread leaf kernel: as many as are resident (two per CU)
    int huge;       // work-groups of the large-leaf sort (one per CU)
    int tilepart;   // work-groups of the tile-local partition (one per CU)
};

// Launchers (all stream-ordered, no host synchronisation).
// flags: SEG_* of the single range, plus kInitLeaf (the range is a leaf of
// the leaf list) or kInitHuge (a leaf of the large-leaf list).
constexpr uint32_t kInitLeaf = 0x100u, kInitHuge = 0x200u;
// Zeroes ctl[0, nblocks) except, with keep_status, the sticky status word
// ctl[kSeedCtl].overflow (later runs of one call keep what earlier runs set).
// The single range [0, n) goes to seg0 (a segment to partition) or, with
// kInitLeaf / kInitHuge, to leaf0 (the leaf list or the large-leaf list).
hipError_t launch_init(LevelCtl* ctl, int nblocks, bool keep_status, SegRec* seg0, LeafRec* leaf0,
                       uint32_t seg0_count, uint64_t n, uint32_t flags, hipStream_t st);
hipError_t launch_leaf(int dtype, const LeafArgs& a, const Grid& g, hipStream_t st);
// Launchers of the key-type kernels (kernels_typed.hip, one instantiation per
// key type): level steps 0 (splitters), 2 (count) and 6 (scatter), and the
// local-sort launches of one pass.
template <int DT> hipError_t typed_level_step(int step, const LevelArgs& a, const Grid& g, hipStream_t st);
template <int DT> hipError_t typed_leaf(const LeafArgs& a, const Grid& g, hipStream_t st);
// Arrays of a segmented sort: above leaf_max to the level-0 queue (counted in
// ctl0->nseg), above leaf_cap to the large-leaf list and the rest to the leaf
// list (both counted in seed->nhuge / seed->nleaf_list).  Arrays whose bounds
// decrease or pass total_keys are skipped and flag kBadOffsets in
// seed->overflow; a device count above nseg_in flags kBadCount.
struct SeedCaps {
    uint32_t segs, leaves, huge;  // entries of the level-0 queue, the leaf list and the large-leaf list
};
hipError_t launch_seed_segments(const uint64_t* offsets, uint32_t nseg_in, const uint32_t* nseg_dev,
                                uint64_t total_keys, SegRec* segs, LevelCtl* ctl0, LevelCtl* seed, LeafRec* leaves,
                                uint32_t leaf_max, uint32_t leaf_cap, LeafRec* huge, SeedCaps caps, hipStream_t st);

// Merge of two sorted runs a[0, na) and b[0, nb) into out[0, na + nb) (out
// overlaps neither run).  `split` holds merge_split_entries(dtype, na + nb)
// words of scratch.
hipError_t launch_merge(int dtype, const void* a, uint64_t na, const void* b, uint64_t nb, void* out, uint64_t* split,
                        hipStream_t st);
uint64_t merge_split_entries(int dtype, uint64_t n);
// The same merge in two parts: the tile boundaries once, then the output
// tiles [tile_begin, tile_end) (tiles of merge_tile_keys(dtype) outputs), so
// that a caller can act on each finished output range.
hipError_t launch_merge_split(int dtype, const void* a, uint64_t na, const void* b, uint64_t nb, uint64_t* split,
                              hipStream_t st);
hipError_t launch_merge_tiles(int dtype, const void* a, uint64_t na, const void* b, uint64_t nb, void* out,
                              const uint64_t* split, uint64_t tile_begin, uint64_t tile_end, hipStream_t st);
uint64_t merge_tile_keys(int dtype);

}  // namespace gqs
