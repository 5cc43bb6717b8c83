// kernels.hip -- gfx950 kernels of the GPU-Quicksort hot path: the chain of one
// partition level and its launch dispatch.
//
// One partition level (the reference's gqsort_kernel launch plus the host
// bookkeeping around it, CL12/QuicksortMain.cpp:332-373 and
// CL12/QuicksortKernels.cl:102-240) is a fixed chain of stream-ordered
// kernels over HBM-resident queues, so no level needs the host:
//
//   k_splitters  per segment: pivot(s).  MODE_CLASSIC takes the median of the
//                first, middle and last key (CL12/QuicksortKernels.cl:218-225);
//                MODE_MULTIPIVOT sorts a stratified sample (wave register
//                sorts, merges across waves through LDS) and keeps up
//                to 255 quantile splitters, giving repeated values their own
//                equality bucket (the generalisation of the reference's
//                "== pivot" class, whose keys are filled in place, :214-216).
//   k_tileplan   prefix sums over segments: tile, count-matrix and bucket bases
//                (the role of the host's blocksize/parent/block loop, :333-354).
//   k_count      persistent over tiles: per-tile bucket histogram in LDS,
//                wave-aggregated through ballot (the reference's per-thread
//                lt/gt counts, :135-145).
//   k_scan_*     exclusive scan of the bucket-major count matrix: every tile's
//                output offset per bucket (replaces the per-parent atomics on
//                sstart/send, :176-187; deterministic and XCD-safe).
//   k_plan       per bucket: size, placement flag, child segments appended to
//                the next level's queue, leaves above one 512-thread work-
//                group's capacity to the large-leaf list (the host's news ->
//                work/done split,
//                :362-371, and CL20's relauncher_kernel, CL20/QuicksortKernels.cl:132-222).
//   k_scatter    persistent over tiles: run starts from the tile's counts in
//                the scanned count matrix, classify again, rank with wave
//                ballots + returning LDS adds that give each key its staging
//                slot, stage the tile in LDS grouped by bucket and store
//                coalesced runs (the scatter pass, :195-203).
//                Keys of equality buckets are not stored at all.
//   equality buckets: the splitter value written over the bucket's range of
//                d, in caller representation, by the work-groups of the next
//                k_leaf_small launch, after the level's scatter (the pivot
//                fill, :214-216, which the reference does serially in the last
//                work-group).
//   k_leaf_*     persistent over buckets: local sort of every bucket of at most
//                leaf_max keys (lqsort_kernel, :267-421), written to d in caller
//                representation.  k_leaf_small: one wave per leaf of at most
//                bitonic_cutoff keys, bitonic network in registers.
//                k_leaf_big: one work-group per larger leaf, software-pipelined
//                (next leaf copied global->LDS asynchronously while the current
//                one is partitioned by interpolation into LDS and its local
//                buckets sorted in registers by odd-even merge networks);
//                512 threads up to 8192 keys (f64: 4096), 1024 threads for
//                the large-leaf list (up to twice that).
//
// Loads that a later wait would serialise are written unconditionally (index
// clamped, value masked), and register networks take their sizes as template
// parameters: both keep clang from waiting for every prefetch load on the
// spot or indexing key registers at run time.
//
// Keys are sorted as unsigned integers: float/double bits go through an
// order-preserving bijection on the first load and back on the final store.
//
// The kernels that depend on the key type (k_splitters, k_count, k_scatter,
// k_leaf_*) and the device helpers live in sortdev.h and are instantiated by
// kernels_typed.hip once per key type; this file holds the others.
#include "sortdev.h"

namespace gqs {

// ------------------------------------------------------------ init / seed
// Sets a bound-hit bit on the level's control block and on the call's sticky
// status word.
__device__ __forceinline__ void flag_overflow(LevelCtl* ctl, uint32_t* status, uint32_t bits) {
    atomicOr(&ctl->overflow, bits);
    atomicOr(status, bits);
}

__global__ void k_init(LevelCtl* ctl, int nblocks, uint32_t keep_status, SegRec* seg0, LeafRec* leaf0,
                       uint32_t seg0_count, uint64_t n, uint32_t flags) {
    constexpr uint32_t kWords = (uint32_t)(sizeof(LevelCtl) / 4);
    const uint32_t words = (uint32_t)nblocks * kWords;
    const uint32_t status_word = (uint32_t)kSeedCtl * kWords + (uint32_t)(offsetof(LevelCtl, overflow) / 4);
    uint32_t* w = reinterpret_cast<uint32_t*>(ctl);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x)
        if (!keep_status || i != status_word) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0 && seg0_count) {
        if (flags & (kInitHuge | kInitLeaf)) {
            leaf0[0] = LeafRec{0, (uint32_t)n, flags & 0xFFu, 0, 0, 0};
            if (flags & kInitHuge) ctl[kSeedCtl].nhuge = 1;  // the single range is a large leaf
            else ctl[kSeedCtl].nleaf_list = 1;               // ... a leaf
        } else {
            seg0[0] = SegRec{0, (uint32_t)n, flags & 0xFFu};
            ctl[0].nseg = 1;
        }
    }
}

// Batched sort: one segment or leaf per caller array (offsets[i]..offsets[i+1]).
// nseg_dev, when given, holds the number of arrays (at most nseg_in, the
// capacity the workspace was sized for) and is read when the kernel runs.
// Arrays to partition go to level 0's queue; arrays sorted directly go to the
// seed block's lists, so that level 0's own large leaves (appended to
// ctl0->nhuge by its plan) do not share a counter with them.
__global__ void k_seed(const uint64_t* offsets, uint32_t nseg_in, const uint32_t* nseg_dev, uint64_t total_keys,
                       SegRec* segs, LevelCtl* ctl0, LevelCtl* seed, LeafRec* leaves, uint32_t leaf_max,
                       uint32_t leaf_cap, LeafRec* huge, SeedCaps caps) {
    uint32_t nseg = nseg_in;
    if (nseg_dev) {
        nseg = *nseg_dev;
        if (nseg > nseg_in) {
            if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&seed->overflow, kBadCount);
            nseg = nseg_in;
        }
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nseg; i += gridDim.x * blockDim.x) {
        const uint64_t lo = offsets[i], hi = offsets[i + 1];
        if (hi < lo || hi > total_keys) {
            atomicOr(&seed->overflow, kBadOffsets);
            continue;
        }
        const uint64_t len = hi - lo;
        if (len < 2) continue;
        // Disjoint arrays within total_keys never exceed the capacities (they
        // are sized from total_keys); the checks keep any other input in bounds.
        uint32_t slot;
        bool ok;
        if (len > leaf_max) {
            slot = atomicAdd(&ctl0->nseg, 1u);
            if ((ok = slot < caps.segs)) segs[slot] = SegRec{lo, (uint32_t)len, SEG_RAW};
        } else {
            const LeafRec r{lo, (uint32_t)len, SEG_RAW, 0, 0, 0};
            if (len > leaf_cap) {
                slot = atomicAdd(&seed->nhuge, 1u);
                if ((ok = slot < caps.huge)) huge[slot] = r;
            } else {
                slot = atomicAdd(&seed->nleaf_list, 1u);
                if ((ok = slot < caps.leaves)) leaves[slot] = r;
            }
        }
        if (!ok) atomicOr(&seed->overflow, kOvfSegs);
    }
}

// ------------------------------------------------------------ k_tileplan
constexpr int kPlanWG = 1024;
__global__ __launch_bounds__(kPlanWG) void k_tileplan(LevelArgs a) {
    __shared__ uint32_t wsum[kPlanWG / 64];
    // A count above the queue's capacity (children the previous level could
    // not store) is clamped to the entries that exist, flagged, and written
    // back below so that every later kernel of the level sees the clamped one.
    const uint32_t nseg_raw = a.ctl->nseg;
    const uint32_t nseg = min(nseg_raw, a.seg_cap);
    uint32_t ct = 0, cm = 0, cb = 0;
    uint64_t keys = 0;
    for (uint32_t base = 0; base < nseg; base += kPlanWG) {
        const uint32_t s = base + threadIdx.x;
        uint32_t nt = 0, nb = 0;
        if (s < nseg) {
            nt = a.plans[s].ntiles;
            nb = a.plans[s].nbuckets;
            keys += a.segs[s].n;
        }
        uint32_t tt, tm, tb;
        // a tile-local level stores one row more per segment: the tile sizes
        const uint32_t rows = nb + (a.tile_local ? 1u : 0u);
        const uint32_t et = block_excl_scan<kPlanWG>(nt, wsum, &tt);
        const uint32_t em = block_excl_scan<kPlanWG>(nt * rows, wsum, &tm);
        const uint32_t eb = block_excl_scan<kPlanWG>(nb, wsum, &tb);
        if (s < nseg) {
            a.plans[s].tile_begin = ct + et;
            a.plans[s].mat_base = cm + em;
            a.plans[s].bucket_base = cb + eb;
        }
        ct += tt;
        cm += tm;
        cb += tb;
    }
    // per-thread key counts summed through LDS atomics (statistics only)
    __shared__ unsigned long long skeys;
    if (threadIdx.x == 0) skeys = 0;
    __syncthreads();
    if (keys) atomicAdd(&skeys, (unsigned long long)keys);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (nseg_raw > nseg) {
            flag_overflow(a.ctl, a.status, kOvfSegs);
            a.ctl->nseg = nseg;
        }
        a.ctl->keys_lo = (uint32_t)skeys;
        a.ctl->keys_hi = (uint32_t)(skeys >> 32);
        if (cm > a.mat_cap || cb > a.bkt_cap) {
            // Capacity is sized from provable bounds; should this ever fire the
            // level is skipped and the call reports GQS_EOVERFLOW.
            flag_overflow(a.ctl, a.status, kOvfMatrix);
            a.ctl->nseg = 0;
            ct = cm = cb = 0;
        }
        a.ctl->total_tiles = ct;
        a.ctl->mat_size = cm;
        a.ctl->total_buckets = cb;
        a.ctl->nchunks = (cm + kScanChunk - 1) / kScanChunk;
    }
}

// ------------------------------------------------------------ scan
// Exclusive scan of the level's count matrix (mat_size entries) in chunks of
// kScanChunk: k_scan_reduce sums each chunk, k_scan_apply adds the sums of the
// chunks before its own and scans it.  Thread t holds entries [16t, 16t + 16)
// of a chunk, moved as four 16-byte loads and stores (the matrix is
// 256-byte aligned and chunks start at multiples of 16 KiB).
constexpr int kScanPer = kScanChunk / 256;
static_assert(kScanPer % 4 == 0, "whole 16-byte vectors per thread");

__device__ __forceinline__ void scan_load(const uint32_t* mat, uint32_t lo, uint32_t size, uint32_t (&v)[kScanPer]) {
    if (lo + kScanPer <= size) {
        const uint4* q = reinterpret_cast<const uint4*>(mat + lo);
#pragma unroll
        for (int j = 0; j < kScanPer / 4; ++j) {
            const uint4 x = q[j];
            v[4 * j] = x.x;
            v[4 * j + 1] = x.y;
            v[4 * j + 2] = x.z;
            v[4 * j + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) v[j] = (lo + j < size) ? mat[lo + j] : 0u;
    }
}

__global__ __launch_bounds__(256) void k_scan_reduce(LevelArgs a) {
    __shared__ uint32_t wsum[4];
    const uint32_t nch = a.ctl->nchunks, size = a.ctl->mat_size;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        uint32_t v[kScanPer];
        scan_load(a.mat, c * kScanChunk + threadIdx.x * kScanPer, size, v);
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) sum += v[j];
        sum = block_sum<256>(sum, wsum);
        if (threadIdx.x == 0) a.part[c] = sum;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_scan_apply(LevelArgs a) {
    __shared__ uint32_t wsum[4];
    const uint32_t nch = a.ctl->nchunks, size = a.ctl->mat_size;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        // sums of chunks [0, c): four independent loads per thread per round
        uint32_t pre = 0;
        for (uint32_t i = threadIdx.x; i < c; i += 4 * 256) {
            uint32_t x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = a.part[min(i + u * 256u, c - 1u)];
#pragma unroll
            for (int u = 0; u < 4; ++u) pre += (i + u * 256u < c) ? x[u] : 0u;
        }
        pre = block_sum<256>(pre, wsum);
        const uint32_t lo = c * kScanChunk + threadIdx.x * kScanPer;
        uint32_t v[kScanPer];
        scan_load(a.mat, lo, size, v);
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) sum += v[j];
        uint32_t tot;
        uint32_t run = pre + block_excl_scan<256>(sum, wsum, &tot);
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            const uint32_t x = v[j];
            v[j] = run;
            run += x;
        }
        if (lo + kScanPer <= size) {
            uint4* q = reinterpret_cast<uint4*>(a.mat + lo);
#pragma unroll
            for (int j = 0; j < kScanPer / 4; ++j) q[j] = make_uint4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
        } else {
#pragma unroll
            for (int j = 0; j < kScanPer; ++j)
                if (lo + j < size) a.mat[lo + j] = v[j];
        }
        __syncthreads();
    }
}

// Splitter idx of the level's splitter table, widened to 64 bits.
__device__ __forceinline__ uint64_t spl_key(const LevelArgs& a, uint32_t idx) {
    return a.key_bytes == 8 ? reinterpret_cast<const uint64_t*>(a.spl)[idx]
                            : (uint64_t)reinterpret_cast<const uint32_t*>(a.spl)[idx];
}

// ------------------------------------------------------------ k_plan
// One work-group per segment, one thread per bucket (nbuckets <= 256).
__global__ __launch_bounds__(256) void k_plan(LevelArgs a) {
    __shared__ uint16_t eqi[kSplStride + 1];  // splitter index + 1 of an equality bucket, else 0
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t fbase;
    const uint32_t nseg = min(a.ctl->nseg, a.seg_cap);
    for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const SegRec seg = a.segs[s];
        const SegPlan p = a.plans[s];
        for (uint32_t i = threadIdx.x; i <= (uint32_t)kSplStride; i += 256) eqi[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)kSplStride; i += 256) {
            const uint16_t m = a.meta[(size_t)s * kSplStride + i];
            if (m >> 8) eqi[(m & 0xFFu) + 1] = (uint16_t)(i + 1);
        }
        __syncthreads();
        const uint32_t m0 = a.mat[p.mat_base];
        const uint32_t b = threadIdx.x;
        uint32_t rel = 0, size = 0, nfill = 0;
        bool eq = false;
        if (b < p.nbuckets) {
            rel = a.mat[p.mat_base + b * p.ntiles] - m0;
            const uint32_t rel1 = (b + 1 < p.nbuckets) ? a.mat[p.mat_base + (b + 1) * p.ntiles] - m0 : seg.n;
            size = rel1 - rel;
            eq = eqi[b] != 0 && size > 0;
            // the scatter wrote the keys to the other buffer
            a.bkt[p.bucket_base + b] = LeafRec{seg.start + rel, size | (eq ? kBucketDirect : 0u),
                                               (seg.flags & SEG_SRC_DN) ? 0u : SEG_SRC_DN, 0, 0, 0};
            if (eq) nfill = (size + kFillChunk - 1) / kFillChunk;
            if (!eq && size > a.leaf_max && !a.is_last) {
                const uint32_t slot = atomicAdd(&a.ctl_next->nseg, 1u);
                if (slot < a.seg_cap) {
                    a.next_segs[slot] = SegRec{seg.start + rel, size, (seg.flags & SEG_SRC_DN) ^ SEG_SRC_DN};
                } else {
                    flag_overflow(a.ctl_next, a.status, kOvfSegs);
                }
            } else if (!eq && size > a.leaf_cap && size <= a.leaf_max) {
                // a leaf too large for one 512-thread work-group: the large-leaf list
                const uint32_t slot = atomicAdd(&a.ctl->nhuge, 1u);
                if (slot < a.huge_cap) {
                    a.huge[slot] = LeafRec{seg.start + rel, size, (seg.flags & SEG_SRC_DN) ^ SEG_SRC_DN, 0, 0, 0};
                } else {
                    flag_overflow(a.ctl, a.status, kOvfHuge);
                }
            }
        }
        uint32_t ftotal;
        const uint32_t fex = block_excl_scan<256>(nfill, wsum, &ftotal);
        if (threadIdx.x == 0 && ftotal) fbase = atomicAdd(&a.ctl->nfill, ftotal);
        __syncthreads();
        if (nfill) {
            const uint64_t key = spl_key(a, (uint32_t)s * kSplStride + (eqi[b] - 1u));
            for (uint32_t c = 0; c < nfill; ++c) {
                const uint32_t slot = fbase + fex + c;
                const uint32_t off = c * kFillChunk;
                if (slot < a.fill_cap)
                    a.fill[slot] = FillTask{seg.start + rel + off, min(kFillChunk, size - off), 0u, key};
                else
                    flag_overflow(a.ctl, a.status, kOvfFill);
            }
        }
        __syncthreads();
    }
}



// ------------------------------------------------------------ k_plan_local
// Plan of a tile-local level (after k_tilepart), one work-group per segment,
// one thread per bucket.  k_tilepart left in the count matrix, for every tile
// t of the segment, the in-tile start S(b, t) of each bucket and, as row nb,
// the tile's size, and added every tile's count of bucket b to bsum; the
// exclusive scan of bsum gives each bucket's place in the sorted segment.
// Per bucket: a gathered leaf record (final position in d, the tiles'
// base in dn, its row of in-tile starts), the large-leaf list for buckets
// above one 512-thread work-group's capacity, fill tasks for equality
// buckets, and a rescue segment (partitioned by the next level from d, where
// the leaf pass first gathers it) for a bucket above leaf_max.
__global__ __launch_bounds__(256) void k_plan_local(LevelArgs a) {
    __shared__ uint16_t eqi[kSplStride + 1];  // splitter index + 1 of an equality bucket, else 0
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t fbase;
    __shared__ uint32_t rsum[kSplStride + 1];  // bucket starts in the segment, then its size
    const uint32_t nseg = min(a.ctl->nseg, a.seg_cap);
    for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const SegRec seg = a.segs[s];
        const SegPlan p = a.plans[s];
        for (uint32_t i = threadIdx.x; i <= (uint32_t)kSplStride; i += 256) eqi[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)kSplStride; i += 256) {
            const uint16_t m = a.meta[(size_t)s * kSplStride + i];
            if (m >> 8) eqi[(m & 0xFFu) + 1] = (uint16_t)(i + 1);
        }
        // bucket sizes (k_tilepart added every tile's counts) and their
        // exclusive scan: where each bucket starts in the sorted segment
        const uint32_t b = threadIdx.x;
        const uint32_t bs = (b < p.nbuckets) ? a.bsum[(size_t)s * kSplStride + b] : 0u;
        uint32_t stot;
        rsum[b] = block_excl_scan<256>(bs, wsum, &stot);
        if (b == 0) rsum[p.nbuckets] = stot;
        __syncthreads();
        uint32_t rel = 0, size = 0, nfill = 0;
        if (b < p.nbuckets) {
            rel = rsum[b];
            size = rsum[b + 1] - rel;
            const bool eq = eqi[b] != 0 && size > 0;
            a.bkt[p.bucket_base + b] = LeafRec{seg.start + rel, size | (eq ? kBucketDirect : 0u), SEG_GATHER,
                                               seg.start, p.mat_base + b * p.ntiles, p.ntiles};
            if (eq) {
                nfill = (size + kFillChunk - 1) / kFillChunk;
            } else if (size > a.leaf_max) {
                // gathered into d by the leaf pass, then partitioned again
                const uint32_t slot = atomicAdd(&a.ctl_next->nseg, 1u);
                if (slot < a.seg_cap) a.next_segs[slot] = SegRec{seg.start + rel, size, 0u};
                else flag_overflow(a.ctl_next, a.status, kOvfSegs);
            } else if (size > a.leaf_cap) {
                const uint32_t slot = atomicAdd(&a.ctl->nhuge, 1u);
                if (slot < a.huge_cap) a.huge[slot] = a.bkt[p.bucket_base + b];
                else flag_overflow(a.ctl, a.status, kOvfHuge);
            }
        }
        uint32_t ftotal;
        const uint32_t fex = block_excl_scan<256>(nfill, wsum, &ftotal);
        if (threadIdx.x == 0 && ftotal) fbase = atomicAdd(&a.ctl->nfill, ftotal);
        __syncthreads();
        if (nfill) {
            const uint64_t key = spl_key(a, (uint32_t)s * kSplStride + (eqi[b] - 1u));
            for (uint32_t c = 0; c < nfill; ++c) {
                const uint32_t slot = fbase + fex + c;
                const uint32_t off = c * kFillChunk;
                if (slot < a.fill_cap)
                    a.fill[slot] = FillTask{seg.start + rel + off, min(kFillChunk, size - off), 0u, key};
                else
                    flag_overflow(a.ctl, a.status, kOvfFill);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ k_plan_gather
// Plan of an in-place tile-local level (tile_local 2, after k_tilepart and
// the scan of its run lengths), one work-group per segment, one thread per
// bucket: every non-empty bucket becomes a gathered segment of the next level
// (SEG_GATHER: its runs in this level's tiles of d, its final place in the
// sorted segment), and an equality bucket becomes fill tasks that the next
// level's leaf pass runs, once that level has read its runs out of d.  The
// fill tasks carry their value: the next level's splitters overwrite this
// level's table.
__global__ __launch_bounds__(256) void k_plan_gather(LevelArgs a) {
    __shared__ uint16_t eqi[kSplStride + 1];  // splitter index + 1 of an equality bucket, else 0
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t fbase;
    __shared__ uint32_t rsum[kSplStride + 1];  // bucket starts in the segment, then its size
    const uint32_t nseg = min(a.ctl->nseg, a.seg_cap);
    for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const SegRec seg = a.segs[s];
        const SegPlan p = a.plans[s];
        for (uint32_t i = threadIdx.x; i <= (uint32_t)kSplStride; i += 256) eqi[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)kSplStride; i += 256) {
            const uint16_t m = a.meta[(size_t)s * kSplStride + i];
            if (m >> 8) eqi[(m & 0xFFu) + 1] = (uint16_t)(i + 1);
        }
        // bucket starts in the segment from the scanned run lengths: row b
        // of the segment begins after the keys of rows 0 .. b - 1 (row
        // nbuckets, all zero, at the segment's size)
        const uint32_t b = threadIdx.x;
        const uint32_t r0 = a.runlen[p.mat_base];
        for (uint32_t i = b; i <= p.nbuckets; i += 256) rsum[i] = a.runlen[p.mat_base + (size_t)i * p.ntiles] - r0;
        __syncthreads();
        uint32_t rel = 0, size = 0, nfill = 0;
        if (b < p.nbuckets) {
            rel = rsum[b];
            size = rsum[b + 1] - rel;
            if (eqi[b] != 0 && size > 0) {
                nfill = (size + kFillChunk - 1) / kFillChunk;
            } else if (size > 0) {
                const uint32_t slot = atomicAdd(&a.ctl_next->nseg, 1u);
                if (slot < a.seg_cap) {
                    a.next_segs[slot] = SegRec{seg.start + rel, size, SEG_GATHER};
                    a.next_gather[slot] = SegGather{seg.start, p.mat_base + b * p.ntiles, p.ntiles};
                } else
                    flag_overflow(a.ctl_next, a.status, kOvfSegs);
            }
        }
        uint32_t ftotal;
        const uint32_t fex = block_excl_scan<256>(nfill, wsum, &ftotal);
        if (threadIdx.x == 0 && ftotal) fbase = atomicAdd(&a.ctl_next->nfill, ftotal);
        __syncthreads();
        if (nfill) {
            const uint64_t key = spl_key(a, (uint32_t)s * kSplStride + (eqi[b] - 1u));
            for (uint32_t c = 0; c < nfill; ++c) {
                const uint32_t slot = fbase + fex + c;
                const uint32_t off = c * kFillChunk;
                if (slot < a.fill_cap)
                    a.fill[slot] = FillTask{seg.start + rel + off, min(kFillChunk, size - off), 0u, key};
                else
                    flag_overflow(a.ctl_next, a.status, kOvfFill);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ launchers
hipError_t launch_init(LevelCtl* ctl, int nblocks, bool keep_status, SegRec* seg0, LeafRec* leaf0,
                       uint32_t seg0_count, uint64_t n, uint32_t flags, hipStream_t st) {
    hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, st, ctl, nblocks, keep_status ? 1u : 0u, seg0, leaf0, seg0_count,
                       n, flags);
    return hipGetLastError();
}

hipError_t launch_seed_segments(const uint64_t* offsets, uint32_t nseg_in, const uint32_t* nseg_dev,
                                uint64_t total_keys, SegRec* segs, LevelCtl* ctl0, LevelCtl* seed, LeafRec* leaves,
                                uint32_t leaf_max, uint32_t leaf_cap, LeafRec* huge, SeedCaps caps, hipStream_t st) {
    const uint32_t blocks = std::max(1u, std::min(1024u, (nseg_in + 255) / 256));
    hipLaunchKernelGGL(k_seed, dim3(blocks), dim3(256), 0, st, offsets, nseg_in, nseg_dev, (unsigned long long)total_keys,
                       segs, ctl0, seed, leaves, leaf_max, leaf_cap, huge, caps);
    return hipGetLastError();
}

hipError_t launch_level_step(int step, int dtype, const LevelArgs& a, const Grid& g, hipStream_t st) {
    switch (step) {
        case 1: hipLaunchKernelGGL(k_tileplan, dim3(1), dim3(kPlanWG), 0, st, a); return hipGetLastError();
        case 3: hipLaunchKernelGGL(k_scan_reduce, dim3(g.scan), dim3(256), 0, st, a); return hipGetLastError();
        case 4: hipLaunchKernelGGL(k_scan_apply, dim3(g.scan), dim3(256), 0, st, a); return hipGetLastError();
        case 5: hipLaunchKernelGGL(k_plan, dim3(g.plan), dim3(256), 0, st, a); return hipGetLastError();
        case 8: hipLaunchKernelGGL(k_plan_local, dim3(g.plan), dim3(256), 0, st, a); return hipGetLastError();
        case 9: hipLaunchKernelGGL(k_plan_gather, dim3(g.plan), dim3(256), 0, st, a); return hipGetLastError();
        case 0: case 2: case 6: case 7: case 10: break;  // typed: splitters, count, scatter, tilepart (2 kinds)
        default: return hipErrorInvalidValue;
    }
    switch (dtype) {
        case DT_U32: return typed_level_step<DT_U32>(step, a, g, st);
        case DT_F32: return typed_level_step<DT_F32>(step, a, g, st);
        case DT_F64: return typed_level_step<DT_F64>(step, a, g, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_leaf(int dtype, const LeafArgs& a, const Grid& g, hipStream_t st) {
    switch (dtype) {
        case DT_U32: return typed_leaf<DT_U32>(a, g, st);
        case DT_F32: return typed_leaf<DT_F32>(a, g, st);
        case DT_F64: return typed_leaf<DT_F64>(a, g, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace gqs
