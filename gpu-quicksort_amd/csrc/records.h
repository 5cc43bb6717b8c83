// records.h -- records and tunables shared by the host dispatcher and the
// device kernels of the GPU-Quicksort hot path.
//
// The reference keeps three records per level (CL12/Quicksort.h:100-145):
// work_record {start, end, pivot, direction}, parent_record {sstart, send,
// oldstart, oldend, blockcount} and block_record {start, end, pivot,
// direction, parent}.  Here one SegRec plays the work_record role (a key range
// still to be partitioned, with the buffer it lives in), SegPlan carries what
// the parent record's counters carried (where each tile's keys go), and the
// block records are implicit: tile t of a level maps to a segment by a binary
// search over SegPlan::tile_begin.  Pivots live in a per-segment splitter table
// instead of in each record so that one segment can have up to 255 splitters.
#pragma once
#include <stdint.h>

namespace gqs {

// Key representation: every dtype is sorted as an unsigned integer key.
enum DType : int { DT_U32 = 0, DT_F32 = 1, DT_F64 = 2 };

// Partition strategy.
//  MODE_MULTIPIVOT: splitters are quantiles of a sorted sample (up to 255 per
//                   segment, fan-out sized so buckets land near leaf_target).
//  MODE_CLASSIC:    one pivot per segment, the median of the first, middle and
//                   last key (CL12/QuicksortKernels.cl:218-225), leaf cutoff
//                   QUICKSORT_BLOCK_SIZE: the reference's own schedule.
enum Mode : int { MODE_MULTIPIVOT = 0, MODE_CLASSIC = 1 };

// SegRec::flags
constexpr uint32_t SEG_SRC_DN = 1u;  // keys live in dn (else in d)
constexpr uint32_t SEG_RAW = 2u;     // keys are still in caller representation
constexpr uint32_t SEG_GATHER = 4u;  // keys are runs of a tile-local level's tiles (a leaf or a segment)

// A key range to partition: [start, start+n) of d (or of dn with SEG_SRC_DN).
// With SEG_GATHER the segment's keys are instead the n keys of one bucket of
// a segment that an in-place tile-local level partitioned (their runs are
// described by the SegGather entry of the same queue slot); `start` is then
// where they go (the next level writes its tiles to dn from there).
struct SegRec {
    uint64_t start;
    uint32_t n;
    uint32_t flags;
};

// Where the keys of a gathered segment are: bucket b of a segment whose tile
// t (kTileKeys keys of d from src + t * kTileKeys) holds the bucket's keys at
// [S(b, t), S(b + 1, t)), S(b, .) being the row src_mat[row ...] of ntiles
// in-tile starts and S(b + 1, .) the row after it; the run of tile t sits at
// P(b, t) = src_pre[row + t] - src_pre[row] in the segment's key order.
struct SegGather {
    uint64_t src;     // first key of the partitioned segment's tiles in d
    uint32_t row;     // count-matrix index of S(b, 0)
    uint32_t ntiles;  // tiles of the partitioned segment
};

// A leaf to sort locally: an entry of the level's bucket table, of the list
// of arrays a sort starts from, or of the large-leaf list.  A contiguous leaf
// is keys [start, start + n) of d (or of dn with SEG_SRC_DN).  A gathered
// leaf (SEG_GATHER) is one bucket b of a segment partitioned by a tile-local
// level: tile t of the segment (kTileKeys keys of dn from src + t * kTileKeys)
// holds the bucket's keys at [S(b, t), S(b + 1, t)), where S(b, .) is the row
// mat[row ...] of ntiles in-tile starts and S(b + 1, .) the row after it; the
// sorted leaf goes to d[start, start + n).
struct LeafRec {
    uint64_t start;   // first key (contiguous) or final position (gathered)
    uint32_t n;       // keys; in the bucket table bit 31 marks an equality bucket
    uint32_t flags;   // SEG_SRC_DN, SEG_RAW, SEG_GATHER
    uint64_t src;     // gathered: first key of the segment's tiles in dn
    uint32_t row;     // gathered: count-matrix index of S(b, 0)
    uint32_t ntiles;  // gathered: tiles of the segment
};

// Per-segment plan of one level, written by the splitter and tile-plan kernels.
struct SegPlan {
    uint32_t tile_begin;   // first tile of this segment in the level's tile space
    uint32_t ntiles;       // ceil(n / tile_keys)
    uint32_t mat_base;     // first entry of this segment in the count matrix
    uint32_t nbuckets;     // buckets of this segment (m + 1 + #equality buckets)
    uint32_t bucket_base;  // first bucket of this segment in the level's bucket space
    uint32_t has_eq;       // 1 when some bucket is an equality bucket (else bucket = search index)
    uint32_t lut_shift;    // lookup table: cell = (key - lut_lo) >> lut_shift
    uint32_t lut_depth;    // search steps after the lookup (max over cells)
    uint64_t lut_lo;       // lowest splitter
};

// Splitter lookup table: kLutCells cells over [u0, u_last]; cell c holds
// #{splitters < u0 + (c << shift)}, the first candidate of the binary search.
constexpr int kLutBits = 10;
constexpr int kLutCells = 1 << kLutBits;

// Per-level control block in HBM.
struct LevelCtl {
    uint32_t nseg;           // segments to partition at this level (appended by the previous level)
    uint32_t total_tiles;    // sum of SegPlan::ntiles
    uint32_t mat_size;       // count-matrix entries
    uint32_t total_buckets;  // sum of SegPlan::nbuckets
    uint32_t nchunks;        // scan chunks over the count matrix
    uint32_t nleaf_list;     // explicit leaf records (level 0 only)
    uint32_t oversize;       // leaves sorted by the unbounded fallback path
    uint32_t overflow;       // capacity overflow flags (should stay 0)
    uint32_t nfill;          // equality-bucket fill tasks of this level
    uint32_t keys_lo;        // keys in this level's segments (low / high word)
    uint32_t keys_hi;
    uint32_t nhuge;          // leaves of this level in the large-leaf list
    uint32_t leaf_next;      // next entry of the bucket table / leaf list for the leaf kernel to take
    uint32_t huge_next;      // next entry of the large-leaf list to take
};

// One chunk of an equality bucket to fill in d with its key value (the
// reference's pivot fill, CL12/QuicksortKernels.cl:214-216, split so that
// many work-groups share a large bucket).
struct FillTask {
    uint64_t start;
    uint32_t n;
    uint32_t pad;
    uint64_t key;  // the value, in key representation
};
constexpr uint32_t kFillChunk = 65536;

constexpr int kMaxSplitters = 255;
constexpr int kSplStride = 256;  // splitter-table stride per segment
constexpr int kMaxLevels = 48;   // level-control blocks allocated per context

// Partition work-group shape.
constexpr int kPartWG = 256;
// Keys per sub-tile: 16 KiB of keys; a tile is kSubPerTile sub-tiles.
template <class K>
struct TileShape {
    static constexpr int kKeysPerThread = 16 * 4 / (int)sizeof(K);  // 16 (u32) / 8 (u64)
    static constexpr int kSubKeys = kKeysPerThread * kPartWG;      // 4096 / 2048
    static constexpr int kSubPerTile = 4;
    static constexpr int kTileKeys = kSubKeys * kSubPerTile;  // 16384 / 8192
};

constexpr int kScanChunk = 4096;   // count-matrix entries per scan chunk
constexpr int kSampleMax = 4096;   // largest splitter sample
constexpr int kLeafWG = 256;       // local-sort work-group size

}  // namespace gqs
