// sortdev.h -- device code of the GPU-Quicksort kernels shared by the
// translation units: wave and work-group helpers, and the kernels that are
// templates over the key type (splitters, count, scatter, local sorts).
// kernels.hip holds the kernels that do not depend on the key type and the
// launch dispatch; kernels_typed.hip instantiates the typed kernels, once per
// key type (GQS_DT), so the three instantiations compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <type_traits>

#include "internal.h"
#include "kernels.h"
#include "keys.h"

namespace gqs {

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Work-group barrier that orders LDS accesses only: global stores and
// asynchronous global->LDS copies issued before it stay in flight, where the
// fence of __syncthreads() would wait for the copies.
__device__ __forceinline__ void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// A value that every lane holds (read from LDS or global memory at a uniform
// address), moved to scalar registers: the compiler cannot tell that such a
// load gives every lane the same value, and would keep it in vector
// registers, live across the whole tile loop.
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni(uint64_t x) {
    return ((uint64_t)uni((uint32_t)(x >> 32)) << 32) | uni((uint32_t)x);
}
__device__ __forceinline__ SegRec uni(const SegRec& r) { return SegRec{uni(r.start), uni(r.n), uni(r.flags)}; }
__device__ __forceinline__ SegPlan uni(const SegPlan& p) {
    SegPlan u;
    u.tile_begin = uni(p.tile_begin);
    u.ntiles = uni(p.ntiles);
    u.mat_base = uni(p.mat_base);
    u.nbuckets = uni(p.nbuckets);
    u.bucket_base = uni(p.bucket_base);
    u.has_eq = uni(p.has_eq);
    u.lut_shift = uni(p.lut_shift);
    u.lut_depth = uni(p.lut_depth);
    u.lut_lo = uni(p.lut_lo);
    return u;
}
// Wait for every global load, store and global->LDS copy this wave issued.
__device__ __forceinline__ void vm_wait_all() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// A load of data no kernel running now writes (tables of an earlier launch),
// through the constant address space: at a wave-uniform address it becomes a
// scalar load, which the wave waits for on its own counter instead of behind
// every vector load and store it has in flight.
template <class T>
__device__ __forceinline__ T ld_const(const T* p) {
    typedef unsigned int w2 __attribute__((ext_vector_type(2)));
    typedef unsigned int w4 __attribute__((ext_vector_type(4)));
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, *(const __attribute__((address_space(4))) uint32_t*)p);
    } else if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, *(const __attribute__((address_space(4))) w2*)p);
    } else {
        static_assert(sizeof(T) == 16, "4, 8 or 16 bytes");
        return __builtin_bit_cast(T, *(const __attribute__((address_space(4))) w4*)p);
    }
}

// Bounds-checked buffer access to [base, base + bytes): loads past the end
// read 0 and stores past it are dropped by the hardware, so a partial tile
// needs no branch around its loads and stores.  Every wave then issues the
// same number of memory instructions on every path, and the compiler can
// wait for a load by count (vmcnt(N)) instead of draining every later store.
struct BufRange {
    __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ BufRange buf_range(const void* base, uint32_t bytes) {
    // dword 3: raw 32-bit data format, no swizzle (the gfx9 untyped-buffer setting)
    return BufRange{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000)};
}
template <class K>
__device__ __forceinline__ K buf_load(const BufRange& b, uint32_t byte_off) {
    if constexpr (sizeof(K) == 4) {
        return (K)__builtin_amdgcn_raw_buffer_load_b32(b.r, (int)byte_off, 0, 0);
    } else {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)byte_off, 0, 0);
        return (K)(((uint64_t)v.y << 32) | v.x);
    }
}
template <class K>
__device__ __forceinline__ void buf_store(K v, const BufRange& b, uint32_t byte_off) {
    if constexpr (sizeof(K) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, b.r, (int)byte_off, 0, 0);
    } else {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        u32x2 w;
        w.x = (uint32_t)v;
        w.y = (uint32_t)((uint64_t)v >> 32);
        __builtin_amdgcn_raw_buffer_store_b64(w, b.r, (int)byte_off, 0, 0);
    }
}

// Inclusive scan over the 64 lanes of a wave with DPP: row_shr 1, 2, 4, 8
// scan each row of 16 lanes (lanes shifted in from outside the row read 0),
// then row_bcast:15 adds row r's last lane to row r + 1 (rows 1, 3) and
// row_bcast:31 adds lane 31 to rows 2 and 3.  Every active lane takes part.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp_add(uint32_t x) {
    return x + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x = dpp_add<0x111>(x);        // row_shr:1
    x = dpp_add<0x112>(x);        // row_shr:2
    x = dpp_add<0x114>(x);        // row_shr:4
    x = dpp_add<0x118>(x);        // row_shr:8
    x = dpp_add<0x142, 0xA>(x);   // row_bcast:15
    x = dpp_add<0x143, 0xC>(x);   // row_bcast:31
    return x;
}

// Exclusive work-group scan of one value per thread (NT threads).
// kTrailing = false leaves out the barrier that protects wsum from the next
// call; callers that pass another barrier before reusing wsum may use it.
template <int NT, bool kLdsOnly = false, bool kTrailing = true>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    // the wave index in a scalar register (uniform in the wave)
    const uint32_t lane = lane_id(), w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t x = wave_incl_scan(v);
    if (lane == 63) wsum[w] = x;
    if constexpr (kLdsOnly) lds_sync();
    else __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        uint32_t t = wsum[i];
        pre += (i < (int)w) ? t : 0u;
        tot += t;
    }
    if constexpr (kTrailing) {
        if constexpr (kLdsOnly) lds_sync();
        else __syncthreads();
    }
    *total = tot;
    return pre + x - v;
}

template <int NT>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* wsum) {
    uint32_t tot;
    block_excl_scan<NT>(v, wsum, &tot);
    return tot;
}

// Add one to hist[b] for every active lane.  A wave whose lanes all carry the
// same bucket (sorted runs, all-equal input) issues one LDS atomic.
// Waves whose keys fall into at most two buckets (sorted or clustered input:
// runs of nearby keys) update the histogram with one atomic per bucket; other
// waves with one atomic per lane.  All mask arithmetic is on wave-uniform
// ballots.
struct Peel {
    uint64_t m0, m1;  // lanes of the first and second leader's bucket
    uint32_t l0, l1;  // leader lanes
    bool two;         // the wave's valid lanes are m0 | m1
};
__device__ __forceinline__ Peel peel2(uint32_t b, bool valid) {
    Peel p{0, 0, 0, 0, false};
    const uint64_t act = __ballot(valid);
    if (act == 0) return p;
    p.l0 = (uint32_t)__builtin_ctzll(act);
    const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)b, p.l0);
    p.m0 = __ballot(valid && b == b0);
    const uint64_t rest = act & ~p.m0;
    if (rest == 0) {
        p.two = true;
        return p;
    }
    p.l1 = (uint32_t)__builtin_ctzll(rest);
    const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)b, p.l1);
    p.m1 = __ballot(valid && b == b1);
    p.two = (p.m0 | p.m1) == act;
    return p;
}
// Whether a wave's keys look clustered: the first key slot of the sub-tile
// falls into at most two buckets.  Random keys skip the peeling.
__device__ __forceinline__ bool clustered(uint32_t b, bool valid) { return peel2(b, valid).two; }

__device__ __forceinline__ void hist_add2(uint32_t* hist, uint32_t b, bool valid, bool try_peel) {
    if (!try_peel) {
        if (valid) atomicAdd(&hist[b], 1u);
        return;
    }
    const Peel p = peel2(b, valid);
    const uint32_t lane = lane_id();
    if (p.two) {
        if (lane == p.l0) atomicAdd(&hist[b], (uint32_t)__popcll(p.m0));
        if (p.m1 && lane == p.l1) atomicAdd(&hist[b], (uint32_t)__popcll(p.m1));
    } else if (valid) {
        atomicAdd(&hist[b], 1u);
    }
}

// Whether every valid key of the wave's N slots falls into one bucket (sorted
// runs, constant input); b0 is that bucket, l0 the first valid lane of slot 0.
template <int N>
__device__ __forceinline__ bool wave_one_bucket(const uint32_t (&b)[N], const bool (&valid)[N], uint32_t& b0,
                                                uint32_t& l0) {
    const uint64_t v0 = __ballot(valid[0]);
    if (v0 == 0) return false;
    l0 = (uint32_t)__builtin_ctzll(v0);
    b0 = (uint32_t)__builtin_amdgcn_readlane((int)b[0], l0);
    bool same = true;
#pragma unroll
    for (int j = 0; j < N; ++j) same = same && (!valid[j] || b[j] == b0);
    return __ballot(!same) == 0;
}

// LDS atomics written as inline assembly, for kernels with global->LDS copies
// in flight.  The compiler makes every LDS atomic it sees wait for all
// outstanding global->LDS copies (it cannot tell which LDS they write); these
// never touch such a buffer, and written this way they do not wait.  The
// returning forms wait for their own results inside the asm statement.
__device__ __forceinline__ uint32_t lds_addr(const uint32_t* p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) uint32_t*)p;
}
__device__ __forceinline__ void lds_add(uint32_t* p, uint32_t v) {
    asm volatile("ds_add_u32 %0, %1" ::"v"(lds_addr(p)), "v"(v) : "memory");
}
__device__ __forceinline__ uint32_t lds_add_rtn(uint32_t* p, uint32_t v) {
    uint32_t r;
    asm volatile("ds_add_rtn_u32 %0, %1, %2\n\ts_waitcnt lgkmcnt(0)" : "=v"(r) : "v"(lds_addr(p)), "v"(v) : "memory");
    return r;
}
// Four returning LDS adds in one asm statement (each lane: four addresses,
// one increment), waited for together.  The outputs are early-clobber: the
// adds complete while later ones still read their inputs.
__device__ __forceinline__ void lds_add_rtn4(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t v,
                                             uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& r3) {
    asm volatile(
        "ds_add_rtn_u32 %0, %4, %8\n\t"
        "ds_add_rtn_u32 %1, %5, %8\n\t"
        "ds_add_rtn_u32 %2, %6, %8\n\t"
        "ds_add_rtn_u32 %3, %7, %8\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3)
        : "v"(a0), "v"(a1), "v"(a2), "v"(a3), "v"(v)
        : "memory");
}
__device__ __forceinline__ void hist_add_asm(uint32_t* hist, uint32_t b) {
    const uint64_t act = __ballot(1);
    const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
    const uint64_t same = __ballot(b == b0);
    if (same == act) {
        if (mbcnt(act) == 0) lds_add(&hist[b0], (uint32_t)__popcll(act));
    } else {
        lds_add(&hist[b], 1u);
    }
}
// hist_rank for N keys per thread at once: every key's atomic is issued
// before any result is waited for.  Invalid keys (valid[j] false) take no
// slot and get rank 0.
template <int N>
__device__ __forceinline__ void hist_rank_n(uint32_t* hist, const uint32_t (&b)[N], const bool (&valid)[N],
                                            uint32_t (&r)[N]) {
    if (!clustered(b[0], valid[0])) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            r[j] = 0;
            if (valid[j]) r[j] = atomicAdd(&hist[b[j]], 1u);
        }
        return;
    }
    {
        // the whole wave in one bucket: one add for all its keys, ranks in
        // slot-major, lane-minor order
        uint32_t b0, l0;
        if (wave_one_bucket<N>(b, valid, b0, l0)) {
            uint64_t vm[N];
            uint32_t before[N], total = 0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                vm[j] = __ballot(valid[j]);
                before[j] = total;
                total += (uint32_t)__popcll(vm[j]);
            }
            uint32_t base = 0;
            if (lane_id() == l0) base = atomicAdd(&hist[b0], total);
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, l0);
#pragma unroll
            for (int j = 0; j < N; ++j) r[j] = valid[j] ? base + before[j] + mbcnt(vm[j]) : 0u;
            return;
        }
    }
    // as hist_add2: a wave of at most two buckets takes one slot range per
    // bucket (rank = range base + lanes below in the same bucket).  All
    // atomics are issued first; the second pass recomputes the (wave-uniform)
    // peel instead of keeping it, and reads the bases from the leader lanes.
    const uint32_t lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const Peel pl = peel2(b[j], valid[j]);
        r[j] = 0;
        if (pl.two) {
            if (lane == pl.l0) r[j] = atomicAdd(&hist[b[j]], (uint32_t)__popcll(pl.m0));
            if (pl.m1 && lane == pl.l1) r[j] = atomicAdd(&hist[b[j]], (uint32_t)__popcll(pl.m1));
        } else if (valid[j]) {
            r[j] = atomicAdd(&hist[b[j]], 1u);
        }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const Peel pl = peel2(b[j], valid[j]);
        if (pl.two) {
            const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)r[j], pl.l0);
            const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)r[j], pl.l1);
            r[j] = ((pl.m0 >> lane) & 1ull) ? r0 + mbcnt(pl.m0) : r1 + mbcnt(pl.m1);
        }
    }
}

// Same, returning each lane's rank among the tile's keys of its bucket.

// Splitter classification: i = #{splitters < key} by a branch-free search over
// 2^depth - 1 sorted entries (padded with the maximum key), then the bucket
// id from the meta table: base[i] plus one when u[i] has an equality bucket
// and key == u[i].
template <class K>
__device__ __forceinline__ uint32_t classify(K key, const K* sK, const uint16_t* sMeta, uint32_t depth) {
    uint32_t i = 0;
    for (uint32_t step = (1u << depth) >> 1; step > 0; step >>= 1) i += (sK[i + step - 1] < key) ? step : 0u;
    const uint32_t meta = sMeta[i];
    return (meta & 0xFFu) + (((meta >> 8) & 1u) & (uint32_t)(key == sK[i]));
}

// classify() for N keys at once with a compile-time search depth: the N
// searches advance step by step together, so their LDS reads overlap.
template <class K, int D, int N>
__device__ __forceinline__ void classify_n_d(const K (&key)[N], uint32_t (&b)[N], const K* sK,
                                             const uint16_t* sMeta) {
    uint32_t i[N];
#pragma unroll
    for (int j = 0; j < N; ++j) i[j] = 0;
#pragma unroll
    for (int st = D - 1; st >= 0; --st) {
        const uint32_t step = 1u << st;
#pragma unroll
        for (int j = 0; j < N; ++j) i[j] += (sK[i[j] + step - 1] < key[j]) ? step : 0u;
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const uint32_t meta = sMeta[i[j]];
        b[j] = (meta & 0xFFu) + (((meta >> 8) & 1u) & (uint32_t)(key[j] == sK[i[j]]));
    }
}
template <class K, int N>
__device__ __forceinline__ void classify_n(const K (&key)[N], uint32_t (&b)[N], const K* sK, const uint16_t* sMeta,
                                           uint32_t depth) {
    switch (depth) {
        case 0: classify_n_d<K, 0, N>(key, b, sK, sMeta); break;
        case 1: classify_n_d<K, 1, N>(key, b, sK, sMeta); break;
        case 2: classify_n_d<K, 2, N>(key, b, sK, sMeta); break;
        case 3: classify_n_d<K, 3, N>(key, b, sK, sMeta); break;
        case 4: classify_n_d<K, 4, N>(key, b, sK, sMeta); break;
        case 5: classify_n_d<K, 5, N>(key, b, sK, sMeta); break;
        case 6: classify_n_d<K, 6, N>(key, b, sK, sMeta); break;
        case 7: classify_n_d<K, 7, N>(key, b, sK, sMeta); break;
        default: classify_n_d<K, 8, N>(key, b, sK, sMeta); break;
    }
}

// Largest s < count with field(s) <= t, for strictly increasing fields.
template <class F>
__device__ __forceinline__ uint32_t upper_find(uint32_t count, uint32_t t, F field) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (field(mid) <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t hash32(uint32_t a, uint32_t b, uint32_t c) {
    uint64_t x = ((uint64_t)a << 32) ^ ((uint64_t)b * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)c << 17) ^ c;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (uint32_t)x;
}

// In-LDS bitonic sort of buf[0..n) with all comparators ascending: the first
// step of each merge pairs i with its mirror i ^ (k-1), the later steps pair
// i with i + h.  Positions >= n act as +infinity, so a pair whose upper index
// is >= n is left alone and no padding is stored.
template <class K>
__device__ void lds_bitonic(K* buf, uint32_t n, uint32_t tid, uint32_t nthreads) {
    if (n < 2) return;
    uint32_t P = 1;
    while (P < n) P <<= 1;
    const uint32_t half = P >> 1;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        const uint32_t hk = k >> 1;
        for (uint32_t q = tid; q < half; q += nthreads) {
            const uint32_t i = ((q & ~(hk - 1)) << 1) | (q & (hk - 1));
            const uint32_t j = i ^ (k - 1);
            if (j < n) {
                K a = buf[i], b = buf[j];
                if (b < a) { buf[i] = b; buf[j] = a; }
            }
        }
        __syncthreads();
        for (uint32_t h = hk >> 1; h >= 1; h >>= 1) {
            for (uint32_t q = tid; q < half; q += nthreads) {
                const uint32_t i = ((q & ~(h - 1)) << 1) | (q & (h - 1));
                const uint32_t j = i + h;
                if (j < n) {
                    K a = buf[i], b = buf[j];
                    if (b < a) { buf[i] = b; buf[j] = a; }
                }
            }
            __syncthreads();
        }
    }
}

// Work-group of the splitter kernel: four sample keys per thread (4096-key
// samples), so the gather, the sample sort and the lookup-table build each
// take a few steps per thread.
constexpr int kSplWG = 1024;
static_assert(kSampleMax % kSplWG == 0, "whole sample slots per thread");

template <class K>
__device__ __forceinline__ K kmin(K a, K b) { return b < a ? b : a; }
template <class K>
__device__ __forceinline__ K kmax(K a, K b) { return b < a ? a : b; }

template <class K, int E>
__device__ __forceinline__ void wave_bitonic(K (&v)[E]);  // below, with the local sort
template <class K, int E, int NB, int J>
__device__ __forceinline__ void wb_half_cleaners(K (&v)[NB][E], uint32_t lane);

// Sorts buf[0, NW*64*E) with the kSplWG threads (NW waves).  Wave w holds
// block w (B = 64*E keys, element i = lane*E + r in v[r]) in its registers
// for the whole sort: it sorts the block there, and in each bitonic merge of
// blocks only the compare distances that reach another block (the mirror
// step, and distances >= B) go through LDS -- one store, a barrier and a load
// of the partner block each -- while every shorter distance runs on the
// registers.  A 4096-key sample thus takes three LDS exchanges.
template <class K, int E>
__device__ __forceinline__ void sample_sort_blocks(K* buf, uint32_t tid) {
    constexpr uint32_t NW = kSplWG / 64;
    constexpr uint32_t B = 64u * E;
    const uint32_t lane = tid & 63u, w = tid >> 6;
    K v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = buf[w * B + lane * E + r];
    wave_bitonic<K, E>(v);
    auto exchange = [&]() {
        __syncthreads();  // every wave has finished reading buf
#pragma unroll
        for (int r = 0; r < E; ++r) buf[w * B + lane * E + r] = v[r];
        __syncthreads();
    };
#pragma unroll
    for (uint32_t k = 2 * B; k <= NW * B; k <<= 1) {
        // mirror step: element g pairs with g ^ (k-1), i.e. local index
        // B-1-i of block w ^ (k/B - 1), held by lane 63-lane at E-1-r
        exchange();
        {
            const K* pb = buf + (w ^ (k / B - 1u)) * B + (63u - lane) * E;
            const bool lower = (w & ((k / B) >> 1)) == 0;
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const K t = pb[E - 1 - r];
                v[r] = lower ? kmin(v[r], t) : kmax(v[r], t);
            }
        }
#pragma unroll
        for (uint32_t h = k >> 2; h >= B; h >>= 1) {
            exchange();
            const K* pb = buf + (w ^ (h / B)) * B + lane * E;
            const bool lower = (w & (h / B)) == 0;
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const K t = pb[r];
                v[r] = lower ? kmin(v[r], t) : kmax(v[r], t);
            }
        }
        wb_half_cleaners<K, E, 1, (int)(B / 2)>(reinterpret_cast<K (&)[1][E]>(v), lane);
    }
    exchange();
}

// Sorts the splitter sample buf[0, S) (S a power of two) with kSplWG threads;
// buf is sorted and visible to the whole work-group on return.
template <class K>
__device__ void sample_sort(K* buf, uint32_t S, uint32_t tid) {
    switch (S / kSplWG) {
        case 1: sample_sort_blocks<K, 1>(buf, tid); break;
        case 2: sample_sort_blocks<K, 2>(buf, tid); break;
        case 4: sample_sort_blocks<K, 4>(buf, tid); break;
        case 8: sample_sort_blocks<K, 8>(buf, tid); break;
        case 16: sample_sort_blocks<K, 16>(buf, tid); break;
        default: lds_bitonic<K>(buf, S, tid, kSplWG); break;
    }
}

// Same network over a global range (unbounded leaves only; one work-group).
template <class K>
__device__ void global_bitonic(K* buf, uint64_t n, uint32_t tid, uint32_t nthreads) {
    if (n < 2) return;
    uint64_t P = 1;
    while (P < n) P <<= 1;
    const uint64_t half = P >> 1;
    for (uint64_t k = 2; k <= P; k <<= 1) {
        for (uint64_t h = k >> 1; h >= 1; h >>= 1) {
            const bool flip = (h == (k >> 1));
            for (uint64_t q = tid; q < half; q += nthreads) {
                const uint64_t i = ((q & ~(h - 1)) << 1) | (q & (h - 1));
                const uint64_t j = flip ? (i ^ (k - 1)) : (i + h);
                if (j < n) {
                    K a = buf[i], b = buf[j];
                    if (b < a) { buf[i] = b; buf[j] = a; }
                }
            }
            __threadfence_block();
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ k_splitters
// Gathered segments (SEG_GATHER): the index in d of the key at position pos
// of the segment.  The run of tile t starts at P(t) = pre[t] - pre[0] in the
// segment's key order; the key is in the last tile t with P(t) <= pos (that
// run is not empty: P(t + 1) > pos, or t is the last tile).
// The run starts P come from pre: the scanned run lengths in global memory,
// or a copy of that row staged in LDS.
template <class K>
__device__ __forceinline__ uint64_t gathered_index(const LevelArgs& a, const SegGather& g, uint32_t pos,
                                                   const uint32_t* pre) {
    const uint32_t base = pre[0];
    uint32_t lo = 0, hi = g.ntiles;  // the tile is in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] - base <= pos) lo = mid;
        else hi = mid;
    }
    return g.src + (uint64_t)lo * TileShape<K>::kTileKeys + a.src_mat[g.row + lo] + (pos - (pre[lo] - base));
}

template <int DT>
__global__ __launch_bounds__(kSplWG) void k_splitters(LevelArgs a) {
    using K = typename KeyOf<DT>::K;
    __shared__ K samp[kSampleMax];
    __shared__ K cand[kSplStride];
    __shared__ uint32_t wsum[kSplWG / 64];
    __shared__ uint16_t lutw[kLutCells + 1];
    // gathered segments: the row of scanned run lengths, staged so that each
    // sample's search reads LDS (rows of at most kPreRow tiles)
    constexpr uint32_t kPreRow = 16384;
    __shared__ uint32_t prow[kPreRow];
    const uint32_t tid = threadIdx.x;
    const uint32_t nseg = min(a.ctl->nseg, a.seg_cap);  // clamped: see k_tileplan
    for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const SegRec seg = a.segs[s];
        const K* src = (seg.flags & SEG_SRC_DN) ? (const K*)a.dn : (const K*)a.d;
        const bool raw = seg.flags & SEG_RAW;
        const uint32_t n = seg.n;
        // key at position pos of the segment (a gathered segment: through its runs)
        const bool gathered = seg.flags & SEG_GATHER;
        const SegGather gs = gathered ? a.gather[s] : SegGather{0, 0, 0};
        const bool staged = gathered && gs.ntiles <= kPreRow;
        if (staged) {
            for (uint32_t i = tid; i < gs.ntiles; i += kSplWG) prow[i] = a.src_pre[gs.row + i];
            __syncthreads();
        }
        // (the staged and the global search are separate calls: one pointer
        // that may be either would be read by flat loads, which wait for every
        // outstanding memory operation)
        auto key_at = [&](uint32_t pos) -> K {
            if (gathered) {
                if (staged) return src[gathered_index<K>(a, gs, pos, prow)];
                return src[gathered_index<K>(a, gs, pos, a.src_pre + gs.row)];
            }
            return load_key<DT>(src, seg.start + pos, raw);
        };
        K* spl = (K*)a.spl + (size_t)s * kSplStride;
        uint16_t* meta = a.meta + (size_t)s * kSplStride;
        if (a.tile_local)
            for (uint32_t i = tid; i < (uint32_t)kSplStride; i += kSplWG) a.bsum[(size_t)s * kSplStride + i] = 0;
        // Candidate splitters cand[0..ncand), sorted, possibly with repeats.
        uint32_t ncand;
        if (a.mode == MODE_CLASSIC) {
            if (tid == 0) {
                const K x1 = key_at(0);
                const K x2 = key_at(n >> 1);
                const K x3 = key_at(n - 1);
                // median-of-3, CL12/Quicksort.h:59-73
                K m;
                if (x1 < x2) m = (x2 < x3) ? x2 : ((x1 < x3) ? x3 : x1);
                else m = (x1 < x3) ? x1 : ((x2 < x3) ? x3 : x2);
                cand[0] = m;
                cand[1] = m;  // a repeated candidate: the pivot gets its equality bucket
            }
            ncand = 2;
        } else {
            // fan-out: ceil(n / leaf_target) buckets (any count up to
            // max_fanout, so that buckets land near leaf_target rather than
            // between leaf_target / 2 and leaf_target).  The last level, which
            // takes what the planned levels left above leaf_max (skewed or
            // duplicate-heavy segments), splits as finely as it can: the most
            // candidates, so that frequent values get equality buckets.
            const uint32_t want = (uint32_t)(((uint64_t)n + a.leaf_target - 1) / a.leaf_target);
            uint32_t k = a.is_last ? a.max_fanout : max(2u, min(want, a.max_fanout));
            uint32_t S = 1;
            while (S < k * a.oversample && S < (uint32_t)kSampleMax) S <<= 1;
            while (S > n) S >>= 1;
            k = max(2u, min(k, S / 2));
            // stratified random sample: all of a thread's loads issue together
            constexpr uint32_t kPer = kSampleMax / kSplWG;
            K sv[kPer];
#pragma unroll
            for (uint32_t e = 0; e < kPer; ++e) {
                const uint32_t j = tid + e * kSplWG;
                sv[e] = K(0);
                if (j < S) {
                    const uint64_t r = hash32(a.seed, (uint32_t)seg.start ^ (uint32_t)(seg.start >> 32), j) % n;
                    const uint64_t pos = ((uint64_t)j * n + r) / S;
                    sv[e] = key_at((uint32_t)pos);
                }
            }
#pragma unroll
            for (uint32_t e = 0; e < kPer; ++e)
                if (tid + e * kSplWG < S) samp[tid + e * kSplWG] = sv[e];
            __syncthreads();
            sample_sort<K>(samp, S, tid);
            ncand = k - 1;
            for (uint32_t i = tid; i < ncand; i += kSplWG) cand[i] = samp[((uint64_t)(i + 1) * S) / k];
        }
        __syncthreads();
        // Unique splitters u[] with an equality flag for values repeated among
        // the candidates (one thread per candidate; ncand <= 255).
        K c = key_max<K>();
        uint32_t first = 0, eq = 0;
        if (tid < ncand) {
            c = cand[tid];
            first = (tid == 0 || cand[tid - 1] != c) ? 1u : 0u;
            eq = (first && tid + 1 < ncand && cand[tid + 1] == c) ? 1u : 0u;
        }
        uint32_t m;
        const uint32_t uidx = block_excl_scan<kSplWG>(first, wsum, &m);
        uint32_t neq;
        // eq flag moves with its unique index: scan eq over unique order.
        const uint32_t eqbefore = block_excl_scan<kSplWG>(eq, wsum, &neq);
        // Table entry i (0 <= i <= 255): splitter u[i] (or max padding) and
        // meta = base[i] | eq[i] << 8 with base[i] = i + #eq among u[0..i).
        // Written by the thread that owns unique index i.
        if (first) {
            spl[uidx] = c;
            meta[uidx] = (uint16_t)((uidx + eqbefore) | (eq << 8));
        }
        const uint32_t nb = m + 1 + neq;
        for (uint32_t i = m + tid; i < (uint32_t)kSplStride; i += kSplWG) {
            spl[i] = key_max<K>();
            meta[i] = (uint16_t)((i == m) ? (m + neq) : (nb - 1));
        }
        // Lookup table over [u0, u_{m-1}] (the unique splitters, staged in samp).
        K* su = samp;
        if (first) su[uidx] = c;
        __syncthreads();
        const K lo = su[0], range = su[m - 1] - su[0];
        uint32_t bitlen = 0;
        while (bitlen < 8 * sizeof(K) && (range >> bitlen) != 0) ++bitlen;
        const uint32_t shift = bitlen > (uint32_t)kLutBits ? bitlen - kLutBits : 0u;
        for (uint32_t cidx = tid; cidx < (uint32_t)kLutCells; cidx += kSplWG) {
            const K off = (K)cidx << shift;
            uint32_t cnt = m;
            if (off <= range) {
                const K v = lo + off;
                uint32_t l = 0, h = m;  // #{u < v}
                while (l < h) {
                    const uint32_t mid = (l + h) >> 1;
                    if (su[mid] < v) l = mid + 1;
                    else h = mid;
                }
                cnt = l;
            }
            lutw[cidx] = (uint16_t)cnt;
        }
        if (tid == 0) lutw[kLutCells] = (uint16_t)m;
        __syncthreads();
        uint32_t wmax = 1;
        for (uint32_t cidx = tid; cidx < (uint32_t)kLutCells; cidx += kSplWG) {
            wmax = max(wmax, (uint32_t)lutw[cidx + 1] - lutw[cidx] + 1u);
            a.lut[(size_t)s * kLutCells + cidx] = (uint8_t)lutw[cidx];
        }
        // block max of the window sizes
        uint32_t wm = wmax;
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) wm = max(wm, (uint32_t)__shfl_xor(wm, dd, 64));
        if (lane_id() == 0) wsum[tid >> 6] = wm;
        __syncthreads();
        if (tid == 0) {
            uint32_t w = 1;
            for (int i = 0; i < kSplWG / 64; ++i) w = max(w, wsum[i]);
            uint32_t ld = 0;
            while ((1u << ld) < w) ++ld;
            SegPlan p;
            p.tile_begin = 0;
            p.ntiles = (uint32_t)(((uint64_t)n + TileShape<K>::kTileKeys - 1) / TileShape<K>::kTileKeys);
            p.mat_base = 0;
            p.nbuckets = nb;
            p.bucket_base = 0;
            p.has_eq = neq ? 1u : 0u;
            p.lut_shift = shift;
            p.lut_depth = ld;
            p.lut_lo = (uint64_t)lo;
            a.plans[s] = p;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ tile helpers
template <class K>
struct TileLds {
    K sK[2 * kSplStride];  // splitters, padded with the maximum key
    uint8_t sLut[kLutCells];
    // 32-bit keys, search depth >= 1: per cell, the splitter that the first
    // search step compares with and the cell's first candidate, read by one
    // 8-byte LDS load instead of two dependent random reads
    uint2 sLut2[sizeof(K) == 4 ? kLutCells : 1];
    uint16_t sMeta[kSplStride];
    uint32_t hist[kSplStride];
    uint32_t seg;
    uint32_t cur;
};

template <class K>
__device__ __forceinline__ void load_splitters(const LevelArgs& a, uint32_t s, TileLds<K>& L) {
    const K* spl = (const K*)a.spl + (size_t)s * kSplStride;
    const uint16_t* meta = a.meta + (size_t)s * kSplStride;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kSplStride; i += blockDim.x) {
        L.sK[i] = spl[i];
        L.sK[i + kSplStride] = key_max<K>();
        L.sMeta[i] = meta[i];
    }
    const uint32_t* lut = reinterpret_cast<const uint32_t*>(a.lut + (size_t)s * kLutCells);
    uint32_t* dl = reinterpret_cast<uint32_t*>(L.sLut);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kLutCells / 4; i += blockDim.x) dl[i] = lut[i];
    if constexpr (sizeof(K) == 4) {
        const uint32_t depth = uni(a.plans[s].lut_depth);
        if (depth >= 1) {     // uniform
            __syncthreads();  // sK and sLut complete
            const uint32_t off = (1u << (depth - 1)) - 1u;
            for (uint32_t c = threadIdx.x; c < (uint32_t)kLutCells; c += blockDim.x) {
                const uint32_t i0 = L.sLut[c];
                L.sLut2[c] = make_uint2((uint32_t)L.sK[i0 + off], i0);
            }
        }
    }
}

// Classification through the lookup table: the table gives the first
// candidate for #{splitters < key}, D more binary-search steps finish it.
template <class K, int D, bool EQ, int N>
__device__ __forceinline__ void classify_lut_d(const K (&key)[N], uint32_t (&b)[N], const TileLds<K>& L, K lo,
                                               uint32_t shift) {
    uint32_t i[N];
    // 32-bit keys: the first step's splitter comes with the candidate
    constexpr bool kPair = sizeof(K) == 4 && D >= 1;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const K off = key[j] - lo;
        const K cell = (key[j] < lo) ? K(0) : min(off >> shift, (K)(kLutCells - 1));
        if constexpr (kPair) {
            const uint2 e = L.sLut2[(uint32_t)cell];
            i[j] = e.y + ((e.x < (uint32_t)key[j]) ? (1u << (D - 1)) : 0u);
        } else {
            i[j] = L.sLut[(uint32_t)cell];
        }
    }
#pragma unroll
    for (int st = kPair ? D - 2 : D - 1; st >= 0; --st) {
        const uint32_t step = 1u << st;
#pragma unroll
        for (int j = 0; j < N; ++j) i[j] += (L.sK[i[j] + step - 1] < key[j]) ? step : 0u;
    }
    if constexpr (EQ) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const uint32_t meta = L.sMeta[i[j]];
            b[j] = (meta & 0xFFu) + (((meta >> 8) & 1u) & (uint32_t)(key[j] == L.sK[i[j]]));
        }
    } else {
        // no equality bucket: the bucket of a key is its search index
#pragma unroll
        for (int j = 0; j < N; ++j) b[j] = i[j];
    }
}
template <class K, int N>
__device__ __forceinline__ void classify_lut(const K (&key)[N], uint32_t (&b)[N], const TileLds<K>& L,
                                             const SegPlan& p) {
    const K lo = (K)p.lut_lo;
    if (p.has_eq) {
        switch (p.lut_depth) {
            case 0: classify_lut_d<K, 0, true, N>(key, b, L, lo, p.lut_shift); break;
            case 1: classify_lut_d<K, 1, true, N>(key, b, L, lo, p.lut_shift); break;
            case 2: classify_lut_d<K, 2, true, N>(key, b, L, lo, p.lut_shift); break;
            case 3: classify_lut_d<K, 3, true, N>(key, b, L, lo, p.lut_shift); break;
            case 4: classify_lut_d<K, 4, true, N>(key, b, L, lo, p.lut_shift); break;
            case 5: classify_lut_d<K, 5, true, N>(key, b, L, lo, p.lut_shift); break;
            case 6: classify_lut_d<K, 6, true, N>(key, b, L, lo, p.lut_shift); break;
            case 7: classify_lut_d<K, 7, true, N>(key, b, L, lo, p.lut_shift); break;
            default: classify_lut_d<K, 8, true, N>(key, b, L, lo, p.lut_shift); break;
        }
    } else {
        switch (p.lut_depth) {
            case 0: classify_lut_d<K, 0, false, N>(key, b, L, lo, p.lut_shift); break;
            case 1: classify_lut_d<K, 1, false, N>(key, b, L, lo, p.lut_shift); break;
            case 2: classify_lut_d<K, 2, false, N>(key, b, L, lo, p.lut_shift); break;
            case 3: classify_lut_d<K, 3, false, N>(key, b, L, lo, p.lut_shift); break;
            case 4: classify_lut_d<K, 4, false, N>(key, b, L, lo, p.lut_shift); break;
            case 5: classify_lut_d<K, 5, false, N>(key, b, L, lo, p.lut_shift); break;
            case 6: classify_lut_d<K, 6, false, N>(key, b, L, lo, p.lut_shift); break;
            case 7: classify_lut_d<K, 7, false, N>(key, b, L, lo, p.lut_shift); break;
            default: classify_lut_d<K, 8, false, N>(key, b, L, lo, p.lut_shift); break;
        }
    }
}

// ------------------------------------------------------------ k_count
template <int DT>
__global__ __launch_bounds__(kPartWG) void k_count(LevelArgs a) {
    using K = typename KeyOf<DT>::K;
    using TS = TileShape<K>;
    __shared__ TileLds<K> L;
    const uint32_t tid = threadIdx.x;
    const uint32_t total = a.ctl->total_tiles;
    const uint32_t nseg = min(a.ctl->nseg, a.seg_cap);  // clamped: see k_tileplan
    uint32_t cur = 0xFFFFFFFFu;
    // Each work-group takes a contiguous run of tiles, so consecutive tiles
    // mostly share a segment (splitters stay in LDS) and thread 0 advances the
    // segment index instead of searching for it.
    const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = min(t0 + per, total);
    uint32_t seg_i = 0, seg_next = 0;
    if (tid == 0 && t0 < t1) {
        seg_i = upper_find(nseg, t0, [&](uint32_t i) { return a.plans[i].tile_begin; });
        seg_next = (seg_i + 1 < nseg) ? a.plans[seg_i + 1].tile_begin : total;
    }
    for (uint32_t t = t0; t < t1; ++t) {
        if (tid == 0) {
            while (t >= seg_next) {
                ++seg_i;
                seg_next = (seg_i + 1 < nseg) ? a.plans[seg_i + 1].tile_begin : total;
            }
            L.seg = seg_i;
        }
        __syncthreads();
        const uint32_t s = uni(L.seg);
        if (s != cur) {
            load_splitters<K>(a, s, L);
            cur = s;
        }
        L.hist[tid] = 0;
        __syncthreads();
        const SegRec seg = uni(a.segs[s]);
        const SegPlan p = uni(a.plans[s]);
        const K* src = (seg.flags & SEG_SRC_DN) ? (const K*)a.dn : (const K*)a.d;
        const bool raw = seg.flags & SEG_RAW;
        const uint32_t tl = t - p.tile_begin;
        const uint32_t lo = tl * (uint32_t)TS::kTileKeys;
        const uint32_t hi = min(lo + (uint32_t)TS::kTileKeys, seg.n);
        const K* base = src + seg.start;
        // Loads are unconditional (index clamped to the tile's last key, the
        // value masked afterwards): loads under a branch cannot be counted,
        // and the compiler would then wait for the next sub-tile's prefetch
        // before classifying the current one.
        // A full tile (every tile but a segment's last) needs no per-key
        // bounds: only its final prefetch, past the tile, is clamped, and
        // those values are never used.
        auto count_tile = [&](auto full_tile) {
            constexpr bool kFull = decltype(full_tile)::value;
            auto load = [&](uint32_t idx) {
                const K v = load_key<DT>(base, min(idx, hi - 1u), raw);
                if constexpr (kFull) return v;
                else return idx < hi ? v : K(0);
            };
            auto valid = [&](uint32_t idx) { return kFull || idx < hi; };
            K keys[TS::kKeysPerThread];
#pragma unroll
            for (int j = 0; j < TS::kKeysPerThread; ++j) keys[j] = load(lo + j * kPartWG + tid);
            for (uint32_t sub = lo; sub < hi; sub += TS::kSubKeys) {
                // prefetch the next sub-tile while this one is classified
                K nxt[TS::kKeysPerThread];
                const uint32_t nsub = sub + TS::kSubKeys;
#pragma unroll
                for (int j = 0; j < TS::kKeysPerThread; ++j) nxt[j] = load(nsub + j * kPartWG + tid);
                uint32_t bk[TS::kKeysPerThread];
                classify_lut<K, TS::kKeysPerThread>(keys, bk, L, p);
                {
                    // (a one-atomic path for waves in one bucket, as the
                    // scatter has, raised this kernel to 138 VGPRs, three
                    // waves per SIMD, and slowed every input)
                    const bool try_peel = clustered(bk[0], valid(sub + tid));
#pragma unroll
                    for (int j = 0; j < TS::kKeysPerThread; ++j)
                        hist_add2(L.hist, bk[j], valid(sub + j * kPartWG + tid), try_peel);
                }
#pragma unroll
                for (int j = 0; j < TS::kKeysPerThread; ++j) keys[j] = nxt[j];
            }
        };
        if (hi - lo == (uint32_t)TS::kTileKeys) count_tile(std::true_type{});
        else count_tile(std::false_type{});
        __syncthreads();
        for (uint32_t b = tid; b < p.nbuckets; b += kPartWG) a.mat[p.mat_base + b * p.ntiles + tl] = L.hist[b];
        __syncthreads();
    }
}

// ------------------------------------------------------------ k_scatter
// Persistent over tiles; every work-group takes a contiguous run of tiles, so
// consecutive tiles mostly share a segment (splitters, segment record and
// plan stay put) and thread 0 advances the segment index instead of searching
// for it.  One tile (kScatterWG * KPT keys) at a time: the tile's per-bucket
// counts are the differences of adjacent entries of the scanned count matrix
// (the count pass saw the same keys in the same tile), so their scan gives
// the tile-local run starts before any key is ranked; the rank counters start
// there and every returning LDS add hands its key the staging slot directly.
// Then the keys are staged grouped by bucket and the runs stored.  The keys
// of the next tile, when it belongs to the same segment and work-group, are
// loaded while the current one is staged and stored.  Barriers per tile: the
// segment/tile-start one, inside the scan, after the run starts (rank
// counters set), after the staging; the next tile's first barrier keeps its
// writes to shared state behind every wave's stores of this one.
constexpr int kScatterWG = 1024;  // threads of a scatter work-group: one tile per pass
template <int DT>
__global__ __launch_bounds__(kScatterWG, 4) void k_scatter(LevelArgs a) {
    using K = typename KeyOf<DT>::K;
    using TS = TileShape<K>;
    constexpr int KPT = TS::kKeysPerThread;
    constexpr uint32_t kTile = (uint32_t)kScatterWG * KPT;  // keys per tile, staged at once
    static_assert(TS::kTileKeys == (int)kTile, "the count pass's tile is one scatter pass");
    static_assert(kTile <= 65536 && kSplStride <= 256, "bucket and slot share one word");
    constexpr bool kEarlyPrefetch = DT == DT_U32;  // keys need no transform
    __shared__ TileLds<K> L;
    __shared__ uint32_t delta[kSplStride];  // output offset - staging offset, per bucket
    __shared__ uint8_t direct[kSplStride];
    __shared__ K stage[kTile];
    __shared__ uint8_t sbk[kTile];
    __shared__ uint32_t wsum[kScatterWG / 64];
    __shared__ uint32_t s_anydirect;
    __shared__ SegRec s_seg;  // the current segment's record and plan
    __shared__ SegPlan s_plan;
    // Scanned count-matrix entries of kMatBlock consecutive tiles plus the one
    // after them, per bucket (row stride kRow, odd against bank conflicts).  A
    // tile's entries lie ntiles apart, one 128-byte line each; loaded tile by
    // tile they were fetched again for every tile (the L2 turns over between
    // two tiles of a work-group), half as many bytes again as the keys.
    // Loaded per block, each line serves kMatBlock tiles.  Entry (b, ntiles)
    // is entry (b + 1, 0), so column j + 1 - column j is the count of bucket
    // b in tile j for every bucket but the last.
    constexpr uint32_t kMatBlock = 32, kRow = kMatBlock + 1;
    __shared__ uint32_t mblk[kSplStride * kRow];
    const uint32_t tid = threadIdx.x;
    const uint32_t total = a.ctl->total_tiles;
    const uint32_t nseg = min(a.ctl->nseg, a.seg_cap);  // clamped: see k_tileplan
    const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = min(t0 + per, total);
    uint32_t seg_i = 0, seg_next = 0;
    if (tid == 0 && t0 < t1) {
        seg_i = upper_find(nseg, t0, [&](uint32_t i) { return a.plans[i].tile_begin; });
        seg_next = (seg_i + 1 < nseg) ? a.plans[seg_i + 1].tile_begin : total;
    }
    uint32_t cur = 0xFFFFFFFFu;
    uint32_t mb_seg = 0xFFFFFFFFu, mb_lo = 0, mb_hi = 0;  // mblk holds tiles [mb_lo, mb_hi) of mb_seg
    uint32_t mb_base = 0;                                  // first entry of mb_seg's matrix
    K keys[KPT];
    bool have_keys = false;  // keys[] already hold this tile's keys
    for (uint32_t t = t0; t < t1; ++t) {
        // tid, opaque to the compiler: values derived from it (LDS addresses)
        // are recomputed per tile instead of being held, or spilled, across
        // the whole loop
        uint32_t ti;
        asm volatile("" : "=v"(ti) : "0"(tid));
        if (ti == 0) {
            while (t >= seg_next) {
                ++seg_i;
                seg_next = (seg_i + 1 < nseg) ? a.plans[seg_i + 1].tile_begin : total;
            }
            L.seg = seg_i;
            s_anydirect = 0;
        }
        __syncthreads();
        const uint32_t s = uni(L.seg);
        if (s != cur) {
            load_splitters<K>(a, s, L);
            if (ti == 0) {
                s_seg = a.segs[s];
                s_plan = a.plans[s];
            }
            cur = s;
            __syncthreads();
        }
        const SegRec seg = uni(s_seg);
        const SegPlan p = uni(s_plan);
        const uint32_t tl = t - p.tile_begin;
        if (s != mb_seg || tl >= mb_hi) {  // uniform; every thread has read the previous block
            mb_seg = s;
            mb_lo = uni(tl);
            mb_hi = uni(min(min(tl + kMatBlock, p.ntiles), tl + (t1 - t)));
            mb_base = uni(a.mat[p.mat_base]);
            // columns 0 .. mb_hi - mb_lo; every slot is loaded from a clamped,
            // valid entry of the segment's matrix (rows past nbuckets, columns
            // past the block and entry (nbuckets - 1, ntiles) are never read),
            // so the loads issue together
            // Four threads per bucket row, column (ti & 3) + 4q.
            static_assert(kScatterWG == 4 * kSplStride, "four threads per matrix row");
            const uint32_t* m = a.mat + p.mat_base;
            const uint32_t last = p.nbuckets * p.ntiles - 1u;
            constexpr int kCols = (int)(kRow + 3) / 4;
            const uint32_t r = ti >> 2, rb = min(r, p.nbuckets - 1u) * p.ntiles + mb_lo;
            uint32_t mv[kCols];
#pragma unroll
            for (int q = 0; q < kCols; ++q) {
                const uint32_t j = min((ti & 3u) + 4u * q, mb_hi - mb_lo);
                mv[q] = m[min(rb + j, last)];
            }
#pragma unroll
            for (int q = 0; q < kCols; ++q) {
                const uint32_t j = (ti & 3u) + 4u * q;
                if (j < kRow) mblk[r * kRow + j] = mv[q];
            }
            __syncthreads();
        }
        // Thread b owns bucket b: its output offset inside the segment, its
        // key count in this tile, and whether it is an equality bucket.
        uint32_t my_cursor = 0, mycnt = 0, bsize = 0;
        if (ti < p.nbuckets) {
            const uint32_t* row = &mblk[ti * kRow + (tl - mb_lo)];
            my_cursor = row[0] - mb_base;
            mycnt = (ti + 1 < p.nbuckets) ? row[1] - row[0] : 0u;  // the last run needs no count
            bsize = a.bkt[p.bucket_base + ti].n;
        }
        const K* src = (seg.flags & SEG_SRC_DN) ? (const K*)a.dn : (const K*)a.d;
        K* dst = ((seg.flags & SEG_SRC_DN) ? (K*)a.d : (K*)a.dn) + seg.start;
        const bool raw = seg.flags & SEG_RAW;
        const uint32_t lo = tl * kTile;
        const uint32_t hi = min(lo + kTile, seg.n);
        const uint32_t cnt = hi - lo;
        const K* base = src + seg.start;
        // Key loads are unconditional (index clamped into the range, value
        // masked afterwards) and the float key transform is applied only
        // when the keys are used: a transform next to a load under a branch
        // made the compiler wait for every prefetch load on the spot.
        if (!have_keys) {
            K raw_keys[KPT];
#pragma unroll
            for (int j = 0; j < KPT; ++j) raw_keys[j] = base[min(lo + j * kScatterWG + ti, hi - 1u)];
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const K v = raw ? to_key<DT>(raw_keys[j]) : raw_keys[j];
                keys[j] = (lo + j * kScatterWG + ti < hi) ? v : K(0);
            }
        }
        // run starts; the rank counters begin at them
        {
            uint32_t tot;
            const uint32_t ls = block_excl_scan<kScatterWG, false, false>(mycnt, wsum, &tot);
            if (ti < p.nbuckets) {
                L.hist[ti] = ls;
                delta[ti] = my_cursor - ls;
                const bool dir = (bsize & kBucketDirect) != 0;
                direct[ti] = dir ? 1 : 0;
                if (dir) s_anydirect = 1;
            }
        }
        __syncthreads();
        const bool anyd = s_anydirect != 0;
        // bucket (< 256) and staging slot (< kTile) of each key, packed as
        // bucket << 16 | slot to keep the registers free for the prefetched keys
        uint32_t br[KPT];
        {
            uint32_t bk[KPT];
            uint32_t rk[KPT];
            bool valid[KPT];
            classify_lut<K, KPT>(keys, bk, L, p);
#pragma unroll
            for (int j = 0; j < KPT; ++j) valid[j] = lo + j * kScatterWG + ti < hi;
            hist_rank_n<KPT>(L.hist, bk, valid, rk);
#pragma unroll
            for (int j = 0; j < KPT; ++j) br[j] = (bk[j] << 16) | rk[j];
        }
        // prefetch: the next tile when it is this work-group's and in the same
        // segment.  uint32 keys are loaded here, before the staging.  Float
        // keys are loaded once the current keys are staged (both sets live at
        // once, with the transform pending, spill), as unconditional clamped
        // loads whose transform and mask wait until the keys are used.
        const bool next_tile = t + 1 < t1 && tl + 1 < p.ntiles;
        const uint32_t nlo = lo + kTile;
        const uint32_t nhi = next_tile ? min(nlo + kTile, seg.n) : 0u;
        K nxt[KPT];
        if constexpr (kEarlyPrefetch) {
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const uint32_t idx = nlo + j * kScatterWG + ti;
                nxt[j] = (idx < nhi) ? base[idx] : K(0);
            }
        }
        // stage by bucket
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            if (lo + j * kScatterWG + ti < hi) {
                const uint32_t pos = br[j] & 0xFFFFu;
                stage[pos] = keys[j];
                sbk[pos] = (uint8_t)(br[j] >> 16);
            }
        }
        if constexpr (!kEarlyPrefetch) {
            if (next_tile) {  // uniform
#pragma unroll
                for (int j = 0; j < KPT; ++j) nxt[j] = base[min(nlo + j * kScatterWG + ti, nhi - 1u)];
            } else {
#pragma unroll
                for (int j = 0; j < KPT; ++j) nxt[j] = K(0);
            }
        }
        lds_sync();  // staged (LDS only: the prefetched keys stay in flight)
        // coalesced stores of the staged runs, W slots per thread at a time:
        // bucket ids, then offsets and keys, each batch of LDS reads issued
        // together (slots past cnt are read but not stored)
        constexpr int W = 8;
#pragma unroll
        for (int u0 = 0; u0 < KPT; u0 += W) {
            if ((uint32_t)u0 * kScatterWG >= cnt) break;
            uint32_t sb[W], dl[W];
            K sv[W];
#pragma unroll
            for (int u = 0; u < W; ++u) sb[u] = sbk[(u0 + u) * kScatterWG + ti];
#pragma unroll
            for (int u = 0; u < W; ++u) {
                dl[u] = delta[sb[u]];
                sv[u] = stage[(u0 + u) * kScatterWG + ti];
            }
            if (anyd) {
#pragma unroll
                for (int u = 0; u < W; ++u)
                    if (direct[sb[u]]) dl[u] = 0xFFFFFFFFu;  // equality buckets: filled after the level
            }
#pragma unroll
            for (int u = 0; u < W; ++u) {
                const uint32_t i = (u0 + u) * kScatterWG + ti;
                if (i < cnt && dl[u] != 0xFFFFFFFFu) dst[dl[u] + i] = sv[u];
            }
        }
        if constexpr (kEarlyPrefetch) {
#pragma unroll
            for (int j = 0; j < KPT; ++j) keys[j] = nxt[j];
        } else {
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const K v = raw ? to_key<DT>(nxt[j]) : nxt[j];
                keys[j] = (nlo + j * kScatterWG + ti < nhi) ? v : K(0);
            }
        }
        have_keys = next_tile;
    }
}

// ------------------------------------------------------------ gathered tiles
// A level whose segments are gathered (SEG_GATHER: buckets of the in-place
// tile-local level before it) reads each of its tiles, kTileKeys consecutive
// keys of the segment's key order, from the runs that bucket left in the
// partitioned tiles: run t is S(b, t) .. S(b + 1, t) of tile t and sits at
// P(t) = pre[t] - pre[0] in the segment's key order.

// The last tile t of gathered segment g with P(t) <= pos, found by one wave
// (64 probes per round); every lane returns it.
__device__ __forceinline__ uint32_t run_find(const LevelArgs& a, const SegGather& g, uint32_t pos) {
    const uint32_t lane = lane_id();
    const uint32_t* pre = a.src_pre + g.row;
    const uint32_t base = pre[0];
    uint32_t lo = 0, len = g.ntiles;  // the tile is in [lo, lo + len); P(lo) <= pos
    while (len > 1) {
        const uint32_t step = (len + 63) >> 6;
        const bool ok = lane * step < len && pre[lo + lane * step] - base <= pos;
        const uint32_t l = 63u - (uint32_t)__clzll(__ballot(ok));  // lane 0 always holds
        lo += l * step;
        len = min(step, len - l * step);
    }
    return lo;
}

// Run table of a gathered tile (one entry per thread): entry i is tile tb + i
// of the partitioned segment; its run, clipped to the tile's key range
// [lo, hi), covers [rpos[i], rpos[i + 1]) of the tile (relative to lo) and
// starts at key rsrc[i] of the partitioned segment's tiles in d.
struct RunTab {
    uint32_t rpos[kScatterWG + 1];
    uint32_t rsrc[kScatterWG];
    uint32_t wle[kScatterWG / 64];  // per wave: table tiles with P(t) <= hi
    uint32_t more;                  // the table ends before hi: another one follows
};
// The gather of one tile: keys [lo, hi) of gathered segment `seg` (whose
// runs g describes), read from the runs of tiles tb on (P(tb) <= lo).  Every
// thread computes the same job from uniform values; thread 0 keeps the next
// one in LDS for the next iteration.
struct GatherJob {
    SegGather g;
    uint32_t n;        // keys of the segment
    uint32_t lo, hi;
    uint32_t tb;
    uint32_t valid;
    uint32_t tile;     // the tile in the level's tile space
    uint32_t seg;      // its segment
    uint32_t seg_end;  // first tile of the next segment
};
__device__ __forceinline__ SegGather uni(const SegGather& g) { return SegGather{uni(g.src), uni(g.row), uni(g.ntiles)}; }
__device__ __forceinline__ GatherJob uni_job(const GatherJob& s) {
    GatherJob J;
    J.valid = uni(s.valid);
    J.g = uni(s.g);
    J.n = uni(s.n);
    J.lo = uni(s.lo);
    J.hi = uni(s.hi);
    J.tb = uni(s.tb);
    J.tile = uni(s.tile);
    J.seg = uni(s.seg);
    J.seg_end = uni(s.seg_end);
    return J;
}

// Thread tid's row entries for tile t of g: S(b, t) and the scanned run length.
__device__ __forceinline__ void run_rows(const LevelArgs& a, const SegGather& g, uint32_t t, uint32_t& s0,
                                         uint32_t& pre) {
    s0 = 0;
    pre = 0;
    if (t < g.ntiles) {
        s0 = a.src_mat[g.row + t];
        pre = a.src_pre[g.row + t];
    }
}

// The run table of job J from tile tb on, one entry per thread (the first
// ntab threads) from the rows it loaded for tile tb + tid (run_rows).
// Complete after a barrier.  No LDS atomics: copies may be in flight.
template <class K>
__device__ __forceinline__ void run_table(RunTab& T, const LevelArgs& a, const GatherJob& J, uint32_t tb,
                                          uint32_t s0, uint32_t pre, uint32_t tid, uint32_t ntab) {
    constexpr uint32_t kTile = TileShape<K>::kTileKeys;
    const uint32_t base = ld_const(a.src_pre + J.g.row);
    const uint32_t t = tb + tid;
    const bool in = t < J.g.ntiles && tid < ntab;
    const uint32_t P = in ? pre - base : J.n;
    const uint32_t c = min(max(P, J.lo), J.hi);
    if (tid < ntab) {
        T.rpos[tid] = c - J.lo;
        T.rsrc[tid] = t * kTile + s0 + (c - P);  // used only where the clipped run is not empty
    }
    const uint32_t le = (uint32_t)__popcll(__ballot(in && P <= J.hi));
    if (lane_id() == 0) T.wle[tid >> 6] = le;
    const uint32_t te = tb + ntab;
    const uint32_t Pe = te < J.g.ntiles ? ld_const(a.src_pre + J.g.row + te) - base : J.n;
    if (tid == 0) {
        T.rpos[ntab] = min(max(Pe, J.lo), J.hi) - J.lo;
        T.more = Pe < J.hi ? 1u : 0u;
    }
}
// Tiles of table T with P(t) <= hi (after the barrier that completes T).
__device__ __forceinline__ uint32_t run_table_le(const RunTab& T) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kScatterWG / 64; ++w) s += T.wle[w];
    return uni(s);
}

// The job of the tile after J's: the next tile of J's segment, whose runs
// start at the last table tile with P(t) <= J.hi (nle of them counted from
// J.tb), or the first tile of the next segment.
template <class K>
__device__ __forceinline__ GatherJob next_job(const LevelArgs& a, const GatherJob& J, uint32_t nle, uint32_t t_end,
                                              uint32_t nseg, uint32_t total) {
    constexpr uint32_t kTile = TileShape<K>::kTileKeys;
    GatherJob N = J;
    N.valid = 0;
    N.g.ntiles = 0;  // (run_rows of an invalid job loads nothing)
    if (!J.valid || J.tile + 1 >= t_end) return N;
    N.g.ntiles = J.g.ntiles;
    N.valid = 1;
    N.tile = J.tile + 1;
    if (N.tile < J.seg_end) {
        N.lo = J.lo + kTile;
        N.tb = J.tb + nle - 1u;
    } else {
        N.seg = J.seg + 1;
        N.g = ld_const(a.gather + N.seg);
        N.n = ld_const(&a.segs[N.seg].n);
        N.seg_end = N.seg + 1 < nseg ? ld_const(&a.plans[N.seg + 1].tile_begin) : total;
        N.lo = 0;
        N.tb = 0;
    }
    N.hi = min(N.lo + kTile, N.n);
    return N;
}

// The keys of table T that fall in this wave's share of the tile, positions
// [w * span, (w + 1) * span) with span = KPT * 64, loaded into the registers
// of their slots: slot j of lane l holds position w * span + j * 64 + l, so
// every load instruction reads 64 consecutive positions (one run, or the
// end of one and the start of the next: coalesced either way).  Each lane
// follows its own positions through the runs with a pointer into T that
// only moves forward.  Slots outside the table's positions are left as they
// are (another table covers them).
template <class K, int KPT>
__device__ __forceinline__ void run_load(K (&v)[KPT], const RunTab& T, const LevelArgs& a, const GatherJob& J,
                                         uint32_t tid, uint32_t ntab) {
    constexpr uint32_t kSpan = (uint32_t)KPT * 64u;
    const uint32_t lane = tid & 63u, w = uni(tid >> 6);
    const uint32_t cnt = J.hi - J.lo;
    const uint32_t xs = max(w * kSpan, uni(T.rpos[0]));
    const uint32_t xe = min(min((w + 1) * kSpan, cnt), uni(T.rpos[ntab]));
    if (xs >= xe) return;
    // the entry holding position xs: the last i < ntab with rpos[i] <= xs
    // (rpos[0] <= xs): 64 probes 16 apart, then 16 single ones
    const uint32_t p1 = lane * 16u;
    const uint32_t l1 = 63u - (uint32_t)__clzll(__ballot(p1 < ntab && T.rpos[p1] <= xs));
    const uint32_t p2 = l1 * 16u + (lane & 15u);
    const uint32_t l2 = 63u - (uint32_t)__clzll(__ballot(lane < 16u && p2 < ntab && T.rpos[p2] <= xs));
    uint32_t i = l1 * 16u + l2;
    uint32_t rp = T.rpos[i], rp1 = T.rpos[i + 1], rs = T.rsrc[i];
    const K* s0 = (const K*)a.d + J.g.src;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const uint32_t pos = w * kSpan + (uint32_t)j * 64u + lane;
        if (pos >= xs && pos < xe) {
            // forward to the run holding pos; pos < xe <= rpos[ntab] stops it
            // before the last entry
            while (pos >= rp1) {
                ++i;
                rp = rp1;
                rp1 = T.rpos[i + 1];
                rs = T.rsrc[i];
            }
            v[j] = s0[rs + (pos - rp)];
        }
    }
}

// ------------------------------------------------------------ k_tilepart
// A tile-local level: each 16 Ki-key tile is partitioned on its own, its
// keys written back grouped by bucket to the tile's own range (no count pass,
// no global scan), its run starts stored as one row of the count matrix and
// its bucket totals added to bsum.  Persistent over tiles like k_scatter, each
// work-group a contiguous run of tiles.  Per tile: classify the keys, rank
// them within their bucket by returning LDS adds, scan the counts into the
// in-tile bucket starts S(b, t), stage the tile grouped by bucket and store it
// over the tile's range (in place where the segment already lives there: the
// whole tile is in registers before any of it is stored).  S(., t) and, as
// row nb, the tile's size go to the count matrix for the level's plan and for
// the readers of the runs.  Equality buckets are stored too (their value is
// filled into d once nothing reads those positions any more).
//   tile_local 1: contiguous segments, tiles into the same range of dn;
//   tile_local 2: contiguous segments of d, tiles in place in d, run lengths
//                 also stored (runlen, laid out as the count matrix) so that
//                 every bucket can be read as one gathered segment;
//   kGather:      gathered segments (SEG_GATHER, the buckets of a tile_local 2
//                 level), tiles into dn from the segment's start.  A tile's
//                 keys are loaded into registers while the previous tile is
//                 partitioned (as the contiguous kinds prefetch theirs),
//                 from a run table built at the top of that iteration out of
//                 rows loaded one iteration earlier still; each wave takes
//                 a contiguous share of the tile's positions (the slot
//                 layout differs from the contiguous kinds', which the
//                 partition does not depend on).
template <int DT, bool kGather>
__global__ __launch_bounds__(kScatterWG, 4) void k_tilepart(LevelArgs a) {
    using K = typename KeyOf<DT>::K;
    using TS = TileShape<K>;
    constexpr int KPT = TS::kKeysPerThread;
    constexpr uint32_t kTile = (uint32_t)kScatterWG * KPT;
    static_assert(TS::kTileKeys == (int)kTile, "one tile per pass of the work-group");
    static_assert(kTile <= 65536 && kSplStride <= 256, "bucket and rank share one word");
    constexpr bool kEarlyPrefetch = DT == DT_U32 && !kGather;  // keys need no transform
    __shared__ TileLds<K> L;
    __shared__ K stage[kTile];
    __shared__ uint32_t wsum[kScatterWG / 64];
    __shared__ SegRec s_seg;
    __shared__ SegPlan s_plan;
    // gathered input: the next tile's run table and the jobs of the next two
    // tiles (slot tile & 1)
    __shared__ std::conditional_t<kGather, RunTab, uint32_t> T;
    __shared__ GatherJob s_job[kGather ? 2 : 1];
    // the counters of odd tiles (even tiles use L.hist): a tile clears the
    // next tile's counters, so no barrier separates the clear from the adds
    __shared__ uint32_t hist_odd[kSplStride];
    // A tile's count-matrix column (its run starts, and in place its run
    // lengths) gathered over 16 tiles and stored a row segment at a time:
    // stored per tile, the column was 257 scattered 4-byte stores (one cache
    // line each, a quarter more write traffic than the keys' own)
    constexpr uint32_t kCols = 16;
    __shared__ uint32_t mcol[kGather ? 1 : 2][kSplStride + 1][kCols];
    const uint32_t tid = threadIdx.x;
    const uint32_t total = a.ctl->total_tiles;
    const uint32_t nseg = min(a.ctl->nseg, a.seg_cap);
    const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = min(t0 + per, total);
    // run-table entries (one per thread; fewer only to drive the multi-table path)
    const uint32_t ntab = kGather ? min(max(a.run_tab, 64u), (uint32_t)kScatterWG) : 0u;
    // every thread tracks the segment of its tile (wave-uniform, scalar loads)
    uint32_t seg_i = 0, seg_next = 0;
    if (t0 < t1) {
        seg_i = upper_find(nseg, t0, [&](uint32_t i) { return ld_const(&a.plans[i].tile_begin); });
        seg_next = (seg_i + 1 < nseg) ? ld_const(&a.plans[seg_i + 1].tile_begin) : total;
    }
    if (tid < (uint32_t)kSplStride) {
        L.hist[tid] = 0;
        hist_odd[tid] = 0;
    }  // (the first tile's segment load ends with a barrier)
    uint32_t rr_s0 = 0, rr_pre = 0;  // gathered input: rows of the table of the tile after next
    K keys[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) keys[j] = K(0);
    if constexpr (kGather) {
        // the first tile's keys, gathered before the loop
        if (t0 < t1) {
            if (tid == 0) {
                GatherJob& J = s_job[t0 & 1];
                J.g = a.gather[seg_i];
                J.n = a.segs[seg_i].n;
                J.lo = (t0 - a.plans[seg_i].tile_begin) * kTile;
                J.hi = min(J.lo + kTile, J.n);
                J.tb = 0;
                J.valid = 1;
                J.tile = t0;
                J.seg = seg_i;
                J.seg_end = seg_next;
            }
            __syncthreads();
            if (uni(s_job[t0 & 1].lo) > 0) {
                if (tid < 64) {
                    const uint32_t tb = run_find(a, uni(s_job[t0 & 1].g), uni(s_job[t0 & 1].lo));
                    if (tid == 0) s_job[t0 & 1].tb = tb;
                }
                __syncthreads();
            }
            const GatherJob J = uni_job(s_job[t0 & 1]);
            uint32_t tb = J.tb, nle = 0;
            for (;;) {
                uint32_t r0, rp;
                run_rows(a, J.g, tb + tid, r0, rp);
                run_table<K>(T, a, J, tb, r0, rp, tid, ntab);
                __syncthreads();
                nle += run_table_le(T);
                run_load<K, KPT>(keys, T, a, J, tid, ntab);
                const bool more = uni(T.more) != 0;
                __syncthreads();  // every wave is done with T
                if (!more) break;
                tb += ntab;
            }
            const GatherJob J1 = next_job<K>(a, J, nle, t1, nseg, total);
            if (tid == 0) s_job[(t0 + 1) & 1] = J1;
            run_rows(a, J1.g, J1.tb + tid, rr_s0, rr_pre);
        }
    }
    uint32_t cur = 0xFFFFFFFFu;
    bool have_keys = false;  // keys[] already hold this tile's keys
    uint32_t btot = 0;       // keys of bucket ti in this work-group's tiles of the segment so far
    for (uint32_t t = t0; t < t1; ++t) {
        // tid, opaque to the compiler (LDS addresses derived from it are
        // recomputed per tile instead of being held across the loop)
        uint32_t ti;
        asm volatile("" : "=v"(ti) : "0"(tid));
        if (!kGather || ti == 0) {
            while (t >= seg_next) {
                ++seg_i;
                seg_next = (seg_i + 1 < nseg) ? ld_const(&a.plans[seg_i + 1].tile_begin) : total;
            }
        }
        if constexpr (kGather) {
            if (ti == 0) L.seg = seg_i;
        }
        // No barrier opens a tile: every wave is past the previous tile's last
        // barrier (staged), after which only the stores read LDS (stage); this
        // tile writes stage after two more barriers, and its counters were
        // cleared by the previous tile (gathered input: one barrier completes
        // the run table built below).  The barriers in the loop are LDS-only:
        // no wave waits for its stores (only for the loads whose registers it
        // uses).
        // (gathered input: measured faster with the previous schedule — a
        // barrier opening the tile, one set of counters cleared behind it,
        // and the barrier that completes the run table built below)
        if constexpr (kGather) lds_sync();
        uint32_t* const H = (kGather || !(t & 1)) ? L.hist : hist_odd;
        uint32_t* const Hn = (kGather || (t & 1)) ? L.hist : hist_odd;
        const uint32_t s = kGather ? uni(L.seg) : seg_i;
        if (s != cur) {
            load_splitters<K>(a, s, L);
            if (ti == 0) {
                s_seg = a.segs[s];
                s_plan = a.plans[s];
            }
            cur = s;
            __syncthreads();
        }
        const SegRec seg = uni(s_seg);
        const SegPlan p = uni(s_plan);
        const uint32_t tl = t - p.tile_begin;
        // the next tile's counters (last read before the staged barrier);
        // gathered input: this tile's, behind the opening barrier
        if (ti < (uint32_t)kSplStride) Hn[ti] = 0;
        const K* src = (seg.flags & SEG_SRC_DN) ? (const K*)a.dn : (const K*)a.d;
        // tile_local 2 partitions in place; the others write dn
        K* dst = (a.tile_local == 2 ? (K*)src : (K*)a.dn) + seg.start;
        const bool raw = seg.flags & SEG_RAW;
        const uint32_t lo = tl * kTile;
        const uint32_t hi = min(lo + kTile, seg.n);
        const uint32_t cnt = hi - lo;
        const K* base = src + seg.start;
        // position in the tile of slot j of this thread (see run_load)
        auto slot_pos = [&](int j) -> uint32_t {
            if constexpr (kGather) return (ti >> 6) * (uint32_t)(KPT * 64) + (uint32_t)j * 64u + (ti & 63u);
            else return (uint32_t)j * kScatterWG + ti;
        };
        if constexpr (kGather) {
            // this tile's keys arrived in keys[] during the previous iteration;
            // the next tile's run table, from the rows loaded then too
            const GatherJob J1 = uni_job(s_job[(t + 1) & 1]);
            if (J1.valid) run_table<K>(T, a, J1, J1.tb, rr_s0, rr_pre, ti, ntab);
        } else if (!have_keys) {
            K raw_keys[KPT];
            const BufRange rin = buf_range(base + lo, cnt * (uint32_t)sizeof(K));
#pragma unroll
            for (int j = 0; j < KPT; ++j) raw_keys[j] = buf_load<K>(rin, (j * kScatterWG + ti) * (uint32_t)sizeof(K));
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const K v = raw ? to_key<DT>(raw_keys[j]) : raw_keys[j];
                keys[j] = (lo + j * kScatterWG + ti < hi) ? v : K(0);
            }
        }
        if constexpr (kGather) {
            lds_sync();  // run table complete
            // the job and rows of the tile after next, requested before the
            // classification (loaded unconditionally: a register assigned on
            // one branch only costs a wait here).  When the next tile's runs
            // need further tables, the tile after next starts from the first
            // table's last tile (P <= hi still holds there).  s_job[t & 1] was
            // last read before the previous tile's staged barrier.
            const GatherJob J1 = uni_job(s_job[(t + 1) & 1]);
            GatherJob J2 = J1;
            J2.valid = 0;
            J2.g.ntiles = 0;
            if (J1.valid) J2 = next_job<K>(a, J1, run_table_le(T), t1, nseg, total);
            if (ti == 0) s_job[t & 1] = J2;
            run_rows(a, J2.g, J2.tb + ti, rr_s0, rr_pre);
        }
        // bucket (< 256) and rank within the bucket (< kTile), packed
        uint32_t br[KPT];
        {
            uint32_t bk[KPT], rk[KPT];
            bool valid[KPT];
            classify_lut<K, KPT>(keys, bk, L, p);
#pragma unroll
            for (int j = 0; j < KPT; ++j) valid[j] = slot_pos(j) < cnt;
            hist_rank_n<KPT>(H, bk, valid, rk);
#pragma unroll
            for (int j = 0; j < KPT; ++j) br[j] = (bk[j] << 16) | rk[j];
        }
        lds_sync();  // every count is in hist
        // in-tile bucket starts: thread b owns bucket b
        {
            const uint32_t mycnt = (ti < p.nbuckets) ? H[ti] : 0u;
            uint32_t tot;
            const uint32_t st = block_excl_scan<kScatterWG, true, false>(mycnt, wsum, &tot);
            if (ti < p.nbuckets) {
                H[ti] = st;
                mcol[0][ti][tl % kCols] = st;
                // in place: the run lengths too (scanned, they give the
                // runs' places in each bucket's key order)
                if constexpr (!kGather) {
                    if (a.tile_local == 2) mcol[kGather ? 0 : 1][ti][tl % kCols] = mycnt;
                }
                btot += mycnt;
            } else if (ti == p.nbuckets) {
                mcol[0][ti][tl % kCols] = cnt;  // row nb: the tile's size
                if constexpr (!kGather) {
                    if (a.tile_local == 2) mcol[kGather ? 0 : 1][ti][tl % kCols] = 0;
                }
            }
        }
        // prefetch: the next tile when it is this work-group's and in the same
        // segment (uint32 keys now, float keys after the staging, as k_scatter)
        const bool next_tile = !kGather && t + 1 < t1 && tl + 1 < p.ntiles;
        const uint32_t nlo = lo + kTile;
        const uint32_t nhi = next_tile ? min(nlo + kTile, seg.n) : 0u;
        K nxt[KPT];
        // the next tile's range (empty without a next tile: every load reads 0)
        const BufRange rnx = buf_range(base + (next_tile ? nlo : 0u), next_tile ? (nhi - nlo) * (uint32_t)sizeof(K) : 0u);
        if constexpr (kEarlyPrefetch) {
#pragma unroll
            for (int j = 0; j < KPT; ++j) nxt[j] = buf_load<K>(rnx, (j * kScatterWG + ti) * (uint32_t)sizeof(K));
        }
        if constexpr (kGather) {
            // the next tile's keys, through the run table built at the top
            const GatherJob J1 = uni_job(s_job[(t + 1) & 1]);
#pragma unroll
            for (int j = 0; j < KPT; ++j) nxt[j] = K(0);
            if (J1.valid) {
                run_load<K, KPT>(nxt, T, a, J1, ti, ntab);
                if (uni(T.more)) {
                    // runs shorter than a table covers (rare): further tables, in turn
                    uint32_t tb = J1.tb;
                    for (;;) {
                        const bool more = uni(T.more) != 0;
                        __syncthreads();  // every wave is done with T
                        if (!more) break;
                        tb += ntab;
                        uint32_t r0, rp;
                        run_rows(a, J1.g, tb + ti, r0, rp);
                        run_table<K>(T, a, J1, tb, r0, rp, ti, ntab);
                        __syncthreads();
                        run_load<K, KPT>(nxt, T, a, J1, ti, ntab);
                    }
                }
            }
        }
        lds_sync();  // bucket starts in hist
#pragma unroll
        for (int j = 0; j < KPT; ++j)
            if (slot_pos(j) < cnt) stage[H[br[j] >> 16] + (br[j] & 0xFFFFu)] = keys[j];
        if constexpr (!kEarlyPrefetch && !kGather) {
#pragma unroll
            for (int j = 0; j < KPT; ++j) nxt[j] = buf_load<K>(rnx, (j * kScatterWG + ti) * (uint32_t)sizeof(K));
        }
        lds_sync();  // the tile is staged
        // the gathered columns out, after every kCols-th tile of the segment,
        // its last tile and this work-group's last: row r's entries for the
        // columns c0 .. tl are consecutive in the count matrix (before the
        // tile's own stores, so the wait for the next tile's keys at the end
        // of the iteration still counts those stores only)
        // bucket totals: one add per bucket for all of this work-group's
        // tiles of the segment (in place, the plan reads them from the
        // scanned run lengths instead)
        if (tl + 1 == p.ntiles || t + 1 == t1) {
            if (a.tile_local != 2 && ti < p.nbuckets && btot) atomicAdd(&a.bsum[(size_t)s * kSplStride + ti], btot);
            btot = 0;
        }
        if (tl % kCols == kCols - 1 || tl + 1 == p.ntiles || t + 1 == t1) {  // uniform
            const uint32_t first = (t0 > p.tile_begin ? t0 - p.tile_begin : 0u);  // this work-group's first column
            const uint32_t c0 = max(tl - tl % kCols, first);
            const uint32_t nrow = p.nbuckets + 1;
            for (uint32_t e = ti; e < nrow * kCols; e += kScatterWG) {
                const uint32_t r = e / kCols, c = tl - tl % kCols + e % kCols;
                if (c >= c0 && c <= tl) {
                    const size_t at = p.mat_base + (size_t)r * p.ntiles + c;
                    a.mat[at] = mcol[0][r][c % kCols];
                    if constexpr (!kGather) {
                        if (a.tile_local == 2) a.runlen[at] = mcol[kGather ? 0 : 1][r][c % kCols];
                    }
                }
            }
        }
        // the tile back, grouped by bucket: coalesced stores of whole tiles
        {
            const BufRange rout = buf_range(dst + lo, cnt * (uint32_t)sizeof(K));
            K sv[KPT];
#pragma unroll
            for (int j = 0; j < KPT; ++j) sv[j] = stage[j * kScatterWG + ti];
#pragma unroll
            for (int j = 0; j < KPT; ++j) buf_store<K>(sv[j], rout, (j * kScatterWG + ti) * (uint32_t)sizeof(K));
        }
        if constexpr (kGather) {
#pragma unroll
            for (int j = 0; j < KPT; ++j) keys[j] = nxt[j];
        } else {
            if constexpr (kEarlyPrefetch) {
#pragma unroll
                for (int j = 0; j < KPT; ++j) keys[j] = nxt[j];
            } else {
#pragma unroll
                for (int j = 0; j < KPT; ++j) {
                    const K v = raw ? to_key<DT>(nxt[j]) : nxt[j];
                    keys[j] = (nlo + j * kScatterWG + ti < nhi) ? v : K(0);
                }
            }
            have_keys = next_tile;
        }
    }
}

// ------------------------------------------------------------ k_leaf
// Local sort of one leaf (the reference's lqsort_kernel,
// CL12/QuicksortKernels.cl:267-421, with sort_threshold/bitonic_sort as its
// tail, :40-97), reshaped for wave64:
//   * leaves of at most bitonic_cutoff (<= 512) keys, k_leaf_small: one wave
//     sorts the leaf in its registers (lane-major bitonic network, cross-lane
//     steps by DPP and swizzles); the four waves of a work-group take four
//     such leaves at once;
//   * larger leaves, k_leaf_big: one interpolation partition of the whole
//     leaf by the work-group between its minimum and maximum (keys held in
//     registers, ranks from returning LDS adds, keys placed by local bucket in
//     LDS), then every thread sorts its local bucket in registers with an
//     odd-even merge network; buckets above the register capacity go to one
//     wave each (bitonic) or, above bitonic_cutoff (rare), to the whole
//     work-group in LDS.
// largest bucket one thread sorts in registers (64 VGPRs of keys at most)
template <class K>
constexpr uint32_t lane_sort_max() { return sizeof(K) == 4 ? 32u : 16u; }

// Batcher's odd-even merge sort network on N (power of two) registers.  The
// merge sizes P and compare distances D are template parameters, so after
// unrolling every register index is a constant (a loop whose trip count
// depends on an outer loop's counter stays a loop, and v[] is then indexed
// at run time through s_set_gpr_idx).
// Pruned to the first M wires (M <= N): when v[M..N) hold the maximum key, a
// compare-exchange touching one of them leaves both wires as they are (the
// smaller value is already on the lower wire, and a maximum never moves
// down), so dropping every such comparator sorts v[0..M) all the same; for
// N = 32 that is 162, 132 and 103 comparators for M = 28, 24 and 20 instead
// of 191.
template <class K, int N, int P, int D, int M>
__device__ __forceinline__ void oem_step(K (&v)[N]) {
    if constexpr (D >= 1) {
#pragma unroll
        for (int j = D % P; j + D < N; j += 2 * D)
#pragma unroll
            for (int i = 0; i < D; ++i)
                if (i + j + D < M && (i + j) / (2 * P) == (i + j + D) / (2 * P)) {
                    const K lo = kmin(v[i + j], v[i + j + D]), hi = kmax(v[i + j], v[i + j + D]);
                    v[i + j] = lo;
                    v[i + j + D] = hi;
                }
        oem_step<K, N, P, D / 2, M>(v);
    }
}
template <class K, int N, int P, int M>
__device__ __forceinline__ void oem_merge(K (&v)[N]) {
    if constexpr (P < N) {
        oem_step<K, N, P, P, M>(v);
        oem_merge<K, N, 2 * P, M>(v);
    }
}
template <class K, int N, int M = N>
__device__ __forceinline__ void oem_network(K (&v)[N]) {
    static_assert(M <= N, "pruned to at most N wires");
    oem_merge<K, N, 1, M>(v);
}

// One thread sorts A[lo, lo + c) (c <= M <= N) in registers, with the
// N-wire network pruned to M wires.
template <class K, int N, int M = N>
__device__ __forceinline__ void lane_sort(K* A, uint32_t lo, uint32_t c) {
    K v[N];
#pragma unroll
    for (int r = 0; r < N; ++r) v[r] = (r < M && (uint32_t)r < c) ? A[lo + r] : key_max<K>();
    oem_network<K, N, M>(v);
#pragma unroll
    for (int r = 0; r < M; ++r)
        if ((uint32_t)r < c) A[lo + r] = v[r];
}


// Value of `x` held by lane (lane ^ lm).  The masks the bitonic network needs
// map to DPP row/quad permutes where one exists (quad_perm for 1, 2, 3;
// row_half_mirror for 7; row_ror:8 for 8; row_mirror for 15; two row_ror
// for 4), to ds_swizzle within 32 lanes for 16 and 31, and to ds_bpermute
// for 32 and 63.  After unrolling, lm is a constant and the switch folds.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t xlane32(uint32_t x, uint32_t lm) {
    switch (lm) {
        case 1: return dpp32<0xB1>(x);   // quad_perm [1,0,3,2]
        case 2: return dpp32<0x4E>(x);   // quad_perm [2,3,0,1]
        case 3: return dpp32<0x1B>(x);   // quad_perm [3,2,1,0]
        case 4: {                        // row_ror:12 gives lane+4, row_ror:4 gives lane-4
            const uint32_t up = dpp32<0x12C>(x), dn = dpp32<0x124>(x);
            return (lane_id() & 4u) ? dn : up;
        }
        case 7: return dpp32<0x141>(x);  // row_half_mirror
        case 8: return dpp32<0x128>(x);  // row_ror:8
        case 15: return dpp32<0x140>(x); // row_mirror
        case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x401F);  // xor 16 within 32
        case 31: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x7C1F);  // xor 31 within 32
        default: return __shfl_xor(x, lm, 64);
    }
}
template <class K>
__device__ __forceinline__ K shfl_xor_k(K v, uint32_t mask) {
    if constexpr (sizeof(K) == 8) {
        const uint32_t lo = xlane32((uint32_t)v, mask);
        const uint32_t hi = xlane32((uint32_t)(v >> 32), mask);
        return ((K)hi << 32) | lo;
    } else {
        return xlane32(v, mask);
    }
}

// Wave-wide min and max: xor distances D, D/2, ..., 1 as template
// parameters, so each shuffle's pattern is a constant.
template <class K, int D>
__device__ __forceinline__ void wave_minmax(K& mn, K& mx) {
    if constexpr (D >= 1) {
        mn = kmin(mn, shfl_xor_k(mn, D));
        mx = kmax(mx, shfl_xor_k(mx, D));
        wave_minmax<K, D / 2>(mn, mx);
    }
}

template <int E> struct Log2 { static constexpr int v = (E <= 1) ? 0 : 1 + Log2<E / 2>::v; };
template <> struct Log2<1> { static constexpr int v = 0; };

// Bitonic network over NB independent problems of 64*E keys each, held
// lane-major (element i = lane*E + r of problem q lives in v[q][r] of lane
// i/E), all comparators ascending (mirror first step of each merge).  One wave
// sorts the NB problems together, so it has NB*E independent compare-exchange
// chains.  Merge sizes and compare distances are template parameters: every
// register index and shuffle pattern is a constant after inlining.
template <class K, int E, int NB, int J>
__device__ __forceinline__ void wb_half_cleaners(K (&v)[NB][E], uint32_t lane) {
    if constexpr (J >= 1) {
        constexpr int LE = Log2<E>::v;
        if constexpr (J < E) {
#pragma unroll
            for (int q = 0; q < NB; ++q)
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    if (!(r & J)) {
                        const K lo = kmin(v[q][r], v[q][r + J]), hi = kmax(v[q][r], v[q][r + J]);
                        v[q][r] = lo;
                        v[q][r + J] = hi;
                    }
                }
        } else {
            constexpr uint32_t lm = (uint32_t)J >> LE;
            const bool upper = (lane & lm) != 0;
#pragma unroll
            for (int q = 0; q < NB; ++q)
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    const K t = shfl_xor_k(v[q][r], lm);
                    v[q][r] = upper ? kmax(v[q][r], t) : kmin(v[q][r], t);
                }
        }
        wb_half_cleaners<K, E, NB, J / 2>(v, lane);
    }
}
template <class K, int E, int NB, int M>
__device__ __forceinline__ void wb_merges(K (&v)[NB][E], uint32_t lane) {
    constexpr int LE = Log2<E>::v;
    if constexpr (M <= 6 + LE) {
        constexpr int k = 1 << M;
        if constexpr (k <= E) {
#pragma unroll
            for (int q = 0; q < NB; ++q)
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    const int p = r ^ (k - 1);
                    if (p > r) {
                        const K lo = kmin(v[q][r], v[q][p]), hi = kmax(v[q][r], v[q][p]);
                        v[q][r] = lo;
                        v[q][p] = hi;
                    }
                }
        } else {
            constexpr uint32_t lm = (uint32_t)(k - 1) >> LE;
            const bool upper = (lane & ((uint32_t)(k >> 1) >> LE)) != 0;
            K t[NB][E];
#pragma unroll
            for (int q = 0; q < NB; ++q)
#pragma unroll
                for (int r = 0; r < E; ++r) t[q][r] = shfl_xor_k(v[q][E - 1 - r], lm);
#pragma unroll
            for (int q = 0; q < NB; ++q)
#pragma unroll
                for (int r = 0; r < E; ++r) v[q][r] = upper ? kmax(v[q][r], t[q][r]) : kmin(v[q][r], t[q][r]);
        }
        wb_half_cleaners<K, E, NB, (k >> 2)>(v, lane);
        wb_merges<K, E, NB, M + 1>(v, lane);
    }
}
template <class K, int E, int NB>
__device__ __forceinline__ void wave_bitonic_nb(K (&v)[NB][E]) {
    wb_merges<K, E, NB, 1>(v, lane_id());
}
template <class K, int E>
__device__ __forceinline__ void wave_bitonic(K (&v)[E]) {
    wave_bitonic_nb<K, E, 1>(reinterpret_cast<K (&)[1][E]>(v));
}

// One wave sorts up to NB sub-buckets of at most 64*E keys each from LDS:
// sub-bucket q is A[lo[q], lo[q]+sz[q]) and goes to out[lo[q] ...].
template <int DT, int E, int NB>
__device__ __forceinline__ void wave_sort_batch(const typename KeyOf<DT>::K* A, typename KeyOf<DT>::K* out,
                                                const uint32_t (&lo)[NB], const uint32_t (&sz)[NB]) {
    using K = typename KeyOf<DT>::K;
    const uint32_t lane = lane_id();
    K v[NB][E];
#pragma unroll
    for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            v[q][r] = (i < sz[q]) ? A[lo[q] + i] : key_max<K>();
        }
    wave_bitonic_nb<K, E, NB>(v);
#pragma unroll
    for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            if (i < sz[q]) out[lo[q] + i] = from_key<DT>(v[q][r]);
        }
}

// One wave sorts n <= 64*E keys read through `ld(i)` and stores key i with
// `st(i, key)`.  With kClamp every lane reads all E slots (index clamped to
// n - 1, value masked afterwards): global reads under a branch, followed by
// ld's key transform, would be waited for one by one.  Without it (LDS reads,
// which need the registers less) slots past n are not read.
template <class K, int E, bool kClamp, class LD, class ST>
__device__ __forceinline__ void wave_sort_n(uint32_t n, LD ld, ST st) {
    if (n == 0) return;
    const uint32_t lane = lane_id();
    K v[E];
    if constexpr (kClamp) {
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = ld(min(lane * E + r, n - 1u));
#pragma unroll
        for (int r = 0; r < E; ++r)
            if (lane * E + r >= n) v[r] = key_max<K>();
    } else {
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            v[r] = (i < n) ? ld(i) : key_max<K>();
        }
    }
    wave_bitonic<K, E>(v);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const uint32_t i = lane * E + r;
        if (i < n) st(i, v[r]);
    }
}

template <class K, bool kClamp = false, class LD, class ST>
__device__ __forceinline__ void wave_sort_any(uint32_t n, LD ld, ST st) {
    if (n <= 64) wave_sort_n<K, 1, kClamp>(n, ld, st);
    else if (n <= 128) wave_sort_n<K, 2, kClamp>(n, ld, st);
    else if (n <= 256) wave_sort_n<K, 4, kClamp>(n, ld, st);
    else wave_sort_n<K, 8, kClamp>(n, ld, st);
}

struct LeafDesc {
    uint64_t start;   // contiguous: first key; gathered: final position in d
    uint64_t src;     // gathered: first key of the segment's tiles in dn
    uint32_t n;
    uint32_t kind;    // 0 none, 1 wave, 2 512-thread work-group, 3 unbounded, 4 large (1024-thread
                      // work-group), 5 an oversize gathered bucket: copied to d in order for the
                      // next level
    uint32_t from_dn;
    uint32_t raw;
    uint32_t gather;  // runs of a tile-local level's tiles (see LeafRec)
    uint32_t row;     // gathered: count-matrix index of S(b, 0)
    uint32_t ntiles;  // gathered: tiles of the segment
};

// Entry idx of the leaf space as 32 raw bytes (a LeafRec of a leaf list or of
// the level's bucket table), kept raw so that its loads can stay in flight
// until decoded.
struct LeafRaw {
    uint4 lo, hi;
};
static_assert(sizeof(LeafRec) == 32, "leaf records are 32 bytes");
__device__ __forceinline__ LeafRaw leaf_rec(const LeafArgs& a, bool list_mode, uint32_t idx, uint32_t count) {
    if (idx >= count) return LeafRaw{make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
    const uint4* q = list_mode ? reinterpret_cast<const uint4*>(a.list) : reinterpret_cast<const uint4*>(a.lv.bkt);
    return LeafRaw{q[2 * (size_t)idx], q[2 * (size_t)idx + 1]};
}
__device__ __forceinline__ LeafDesc leaf_decode(const LeafArgs& a, bool list_mode, const LeafRaw& r) {
    LeafDesc ds{0, 0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t n = r.lo.z, flags = r.lo.w;
    const bool gather = (flags & SEG_GATHER) != 0;
    // a level's bucket table: equality buckets are filled, buckets above
    // leaf_max of a contiguous level are the next level's segments
    const bool take = list_mode || (!(n & kBucketDirect) && (n <= a.leaf_max || a.is_last || gather));
    if (take && n >= 1) {
        ds.start = (uint64_t)r.lo.x | ((uint64_t)r.lo.y << 32);
        ds.n = n;
        ds.from_dn = (flags & SEG_SRC_DN) ? 1u : 0u;
        ds.raw = (flags & SEG_RAW) ? 1u : 0u;
        ds.gather = gather ? 1u : 0u;
        ds.src = (uint64_t)r.hi.x | ((uint64_t)r.hi.y << 32);
        ds.row = r.hi.z;
        ds.ntiles = r.hi.w;
        // a gathered bucket above leaf_max is the next level's segment (the
        // leaf pass only copies it together), even below leaf_cap
        if (gather) ds.kind = n > a.leaf_max ? 5u : (n <= a.leaf_cap ? 2u : 4u);
        else ds.kind = (n <= a.bitonic_cutoff) ? 1u : (n <= a.leaf_cap ? 2u : (n <= a.leaf_max ? 4u : 3u));
    }
    return ds;
}
__device__ __forceinline__ LeafDesc uniform_desc(const LeafDesc& d) {
    LeafDesc u;
    u.start = uni(d.start);
    u.src = uni(d.src);
    u.n = uni(d.n);
    u.kind = uni(d.kind);
    u.from_dn = uni(d.from_dn);
    u.raw = uni(d.raw);
    u.gather = uni(d.gather);
    u.row = uni(d.row);
    u.ntiles = uni(d.ntiles);
    return u;
}

// Rows of a gathered leaf staged in LDS ahead of its copy: S(b, t) for the
// first kRowsLds tiles, then S(b + 1, t).  A leaf of more tiles reads its rows
// from global memory while it is copied, which waits for every outstanding
// memory operation: twice the expected tile count of a segment (1 Mi keys of
// 16 Ki-key tiles at 2^28 keys) keeps that off the common path.
constexpr uint32_t kRowsLds = 128;

// Asynchronous copy of a gathered leaf into LDS (global_load_lds, no
// registers): run t of the leaf is S(b, t) .. S(b+1, t) of tile t, the rows
// rowb[0, ntiles) and rowb[ntiles, 2 * ntiles), read from rows_lds (staged
// there earlier) when `staged` and ntiles <= kRowsLds.  rows_lds always points
// into LDS (also when not staged): a pointer that might be null or global
// makes the compiler read it with flat loads, and a flat load waits for every
// outstanding memory operation (the previous leaf's stores, the copies in
// flight).  Every wave reads the rows of 64 tiles at a time, scans the run
// lengths itself (DPP) and copies the runs of the tiles t = wave (mod waves),
// 64 dwords per instruction, to their offsets in dst.  Complete after
// vm_wait_all() in every wave and a barrier.
template <class K, int NT>
__device__ __forceinline__ void gather_to_lds_async(K* dst, const K* base, const uint32_t* rowb, uint32_t ntiles,
                                                    const uint32_t* rows_lds, bool staged, uint32_t tid) {
    constexpr uint32_t NW = NT / 64, DW = sizeof(K) / 4;
    constexpr uint32_t kTile = TileShape<K>::kTileKeys;
    const uint32_t lane = tid & 63u, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t* s = reinterpret_cast<const uint32_t*>(base);
    uint32_t* p = reinterpret_cast<uint32_t*>(dst);
    uint32_t off = 0;  // keys of the chunks before this one
    for (uint32_t c0 = 0; c0 < ntiles; c0 += 64) {
        const uint32_t t = c0 + lane;
        uint32_t s0 = 0, len = 0;
        if (t < ntiles) {
            if (staged && ntiles <= kRowsLds) {
                s0 = rows_lds[t];
                len = rows_lds[kRowsLds + t] - s0;
            } else {
                s0 = rowb[t];
                len = rowb[ntiles + t] - s0;
            }
        }
        const uint32_t incl = wave_incl_scan(len);
        const uint32_t excl = incl - len;
        for (uint32_t l = w; l < 64u && c0 + l < ntiles; l += NW) {
            const uint32_t rs0 = (uint32_t)__builtin_amdgcn_readlane((int)s0, (int)l);
            const uint32_t ndw = (uint32_t)__builtin_amdgcn_readlane((int)len, (int)l) * DW;
            const uint32_t roff = (uint32_t)__builtin_amdgcn_readlane((int)excl, (int)l);
            const uint32_t* src = s + ((size_t)(c0 + l) * kTile + rs0) * DW;
            uint32_t* d = p + (size_t)(off + roff) * DW;
            for (uint32_t q = 0; q < ndw; q += 64)
                if (q + lane < ndw)
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + q + lane),
                                                     (__attribute__((address_space(3))) void*)(d + q), 4, 0, 0);
        }
        off += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
}

// Waves 0 and 1 stage the two rows of a gathered leaf of at most kRowsLds
// tiles in rows_lds with asynchronous copies (complete after vm_wait_all() in
// those waves and a barrier).
__device__ __forceinline__ void stage_rows_async(uint32_t* rows_lds, const uint32_t* rowb, uint32_t ntiles,
                                                 uint32_t tid) {
    static_assert(kRowsLds == 128, "two waves, one tile per lane");
    if (tid < kRowsLds && ntiles <= kRowsLds) {
        const uint32_t w = __builtin_amdgcn_readfirstlane(tid >> 6);
        if (w * 64u < ntiles && tid < ntiles) {
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(rowb + tid),
                                             (__attribute__((address_space(3))) void*)(rows_lds + w * 64u), 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(rowb + ntiles + tid),
                                             (__attribute__((address_space(3))) void*)(rows_lds + kRowsLds + w * 64u),
                                             4, 0, 0);
        }
    }
}

// The same runs copied in order to out[0, n) in global memory (an oversize
// gathered bucket, partitioned again by the next level); one work-group, run
// by run.
template <class K, int NT>
__device__ void gather_copy_global(K* out, const K* base, const uint32_t* rowb, uint32_t ntiles, uint32_t tid) {
    constexpr uint32_t kTile = TileShape<K>::kTileKeys;
    uint32_t off = 0;
    for (uint32_t t = 0; t < ntiles; ++t) {
        const uint32_t s0 = uni(rowb[t]), s1 = uni(rowb[ntiles + t]);
        const K* src = base + (size_t)t * kTile + s0;
        for (uint32_t i = tid; i < s1 - s0; i += NT) out[off + i] = src[i];
        off += s1 - s0;
    }
}

// Leaves of at most bitonic_cutoff keys: one per wave, four per work-group of
// kLeafWG threads, each sorted in its wave's registers.  After a partition
// level the same work-groups then write the level's equality buckets: the
// splitter value, in caller representation, over each fill task's range of d.
// This launch follows the level's scatter, so no tile of the level still
// reads those positions, and no leaf or later level touches them.
template <int DT>
__global__ __launch_bounds__(kLeafWG) void k_leaf_small(LeafArgs a) {
    using K = typename KeyOf<DT>::K;
    __shared__ LeafDesc desc[4];
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    const bool list_mode = a.list != nullptr;
    const uint32_t count = list_mode ? min(*a.list_count, a.list_cap) : a.lv.ctl->total_buckets;
    K* dout = (K*)a.d;
    for (uint32_t g = blockIdx.x; g * 4 < count; g += gridDim.x) {
        if (tid < 4) {
            LeafDesc ds = leaf_decode(a, list_mode, leaf_rec(a, list_mode, g * 4 + tid, count));
            if (ds.kind != 1) ds.kind = 0;  // larger leaves: k_leaf_big
            desc[tid] = ds;
        }
        __syncthreads();
        const LeafDesc ds = desc[wave];
        if (ds.kind == 1) {
            const K* src = (ds.from_dn ? (const K*)a.dn : (const K*)a.d) + ds.start;
            K* out = dout + ds.start;
            const bool raw = ds.raw;
            wave_sort_any<K, true>(
                ds.n, [&](uint32_t i) { return load_key<DT>(src, i, raw); },
                [&](uint32_t i, K v) { out[i] = from_key<DT>(v); });
        }
        __syncthreads();  // desc[] is rewritten by the next iteration
    }
    if (!list_mode) {
        const uint32_t nt = min(a.lv.ctl->nfill, a.lv.fill_cap);
        for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
            const FillTask f = a.lv.fill[t];
            const K v = from_key<DT>((K)f.key);
            for (uint32_t i = tid; i < f.n; i += kLeafWG) dout[f.start + i] = v;
        }
    }
}

// Asynchronous copy of n keys from global memory to LDS without passing
// through registers (global_load_lds): wave w moves dwords [64w, 64w + 64) of
// every NT-dword stripe.  Complete after vm_wait_all() in every wave and a
// barrier.
template <class K, int NT>
__device__ __forceinline__ void copy_to_lds_async(K* dst, const K* src, uint32_t n, uint32_t tid) {
    const uint32_t ndw = n * (uint32_t)(sizeof(K) / 4);
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* p = reinterpret_cast<uint32_t*>(dst);
    const uint32_t lane = tid & 63u;
    for (uint32_t base = __builtin_amdgcn_readfirstlane(tid >> 6) * 64u; base < ndw; base += NT) {
        if (base + lane < ndw)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(s + base + lane),
                                             (__attribute__((address_space(3))) void*)(p + base), 4, 0, 0);
    }
}

// Leaves above bitonic_cutoff keys, one per work-group of NTX threads,
// software-pipelined over the work-group's leaves g_i = g0 + i * G:
// while leaf g_i is sorted (keys in registers, LDS buffer A), the keys of
// g_{i+1} stream into LDS buffer P, the record of g_{i+2} is in flight, and
// the sorted g_{i-1} is being written out from A.  Barriers inside the
// iteration order LDS only (lds_sync), so none of that traffic is drained
// before the next iteration starts.
//
// Sorting one leaf: one interpolation partition into about n/sort_threshold
// local buckets between the leaf's minimum and maximum (monotone in the key,
// one bucket per thread), keys placed by bucket in A, then every thread sorts
// its bucket in registers with an odd-even merge network; buckets above the
// register capacity (skewed or duplicate-heavy leaves) go to one wave each,
// or to the work-group in LDS.
//
// One kernel in two sizes, 16 (double: 8) keys per thread and one local
// bucket per thread: NTX = 512 for leaves up to leaf_cap (kind 2; it also
// takes the unbounded fallback, kind 3, and the copies of oversize gathered
// buckets, kind 5); NTX = 1024 for the leaves between leaf_cap and leaf_max
// (kind 4, read from the large-leaf list), so that buckets slightly above the
// 512-thread capacity need no further partition level.  (A 256-thread size
// for leaves up to leaf_cap / 2, four work-groups per CU, was measured slower
// at every leaf size.)
template <int DT, int NTX>
__global__ __launch_bounds__(NTX, 4) void k_leaf_big(LeafArgs a) {
    static_assert(NTX == 512 || NTX == 1024, "two sizes");
    using K = typename KeyOf<DT>::K;
    constexpr int NT = NTX;
    constexpr int NW = NT / 64;
    constexpr uint32_t CAP = ((sizeof(K) == 4) ? 16u : 8u) * NTX;  // largest leaf
    constexpr int KPT = (int)CAP / NT;                             // keys per thread
    constexpr int LL = NT;  // local buckets (one per thread)
    constexpr uint32_t kMine = NTX == 1024 ? 4u : 2u;  // the leaf kind this kernel sorts
    // The small tables lead their LDS block so that their addresses stay
    // below 64 KiB, within reach of the immediate offset of LDS instructions.
    // P is an LDS variable of its own: the compiler then knows that plain LDS
    // loads and stores of the other variables do not touch what the
    // asynchronous copies write, and places no wait for the copies before
    // them (LDS atomics it waits for regardless; those here are inline asm).
    struct Lds {
        LeafDesc sd[4];
        uint32_t recs[2][8];             // raw leaf records, fetched two iterations ahead
        uint32_t rows[2][2 * kRowsLds];  // rows of gathered leaves, staged one iteration ahead
        uint32_t hist[LL + 1];
        uint32_t bst[LL + 1];
        uint32_t s_nbig;
        uint32_t idx[4];    // leaf index of iteration i in idx[i % 4]
        uint32_t junk[64];  // targets of the cursor adds of lanes without a key, one per lane
        uint32_t scanw[NW];
        K wmin[NW], wmax[NW];
        uint16_t bigs16[LL];
        alignas(16) K A[CAP];  // the leaf being sorted, written out next iteration
    };
    __shared__ Lds S;
    __shared__ K P[CAP];  // raw keys of the next leaf, arriving asynchronously
    LeafDesc* sd = S.sd;
    uint32_t* hist = S.hist;
    uint32_t* bst = S.bst;
    uint32_t& s_nbig = S.s_nbig;
    uint32_t* scanw = S.scanw;
    K* wmin = S.wmin;
    K* wmax = S.wmax;
    uint16_t* bigs16 = S.bigs16;
    K* A = S.A;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const bool list_mode = a.list != nullptr;
    const uint32_t count = list_mode ? min(*a.list_count, a.list_cap) : a.lv.ctl->total_buckets;
    K* dout = (K*)a.d;
    auto big_only = [](LeafDesc ds) {
        // kinds sorted elsewhere: 1 (the wave kernel), 4 or 2 (the other
        // instantiation); the unbounded fallback (3) and the copy of an
        // oversize gathered bucket (5) belong to the 512-thread one
        if (ds.kind != kMine && (NTX != 512 || (ds.kind != 3u && ds.kind != 5u))) ds.kind = 0;
        return ds;
    };
    // Leaf records travel to LDS asynchronously (wave 0, one dword per lane),
    // so no register holds a record in flight.
    auto fetch = [&](uint32_t idx, uint32_t slot) {
        if (tid < 8 && idx < count) {
            // a 32-bit offset from the scalar base, so that no 64-bit lane
            // address stays live across the loop (one that did was spilled,
            // and its reload waited for every copy in flight)
            const uint32_t* base =
                list_mode ? reinterpret_cast<const uint32_t*>(a.list) : reinterpret_cast<const uint32_t*>(a.lv.bkt);
            const uint32_t o = idx * 8u + tid;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(base + o),
                                             (__attribute__((address_space(3))) void*)(S.recs[slot]), 4, 0, 0);
        }
    };
    auto decode = [&](uint32_t idx, uint32_t slot) {  // thread 0, after its wave waited for the fetch
        const uint32_t* r = S.recs[slot];
        const LeafRaw raw{make_uint4(r[0], r[1], r[2], r[3]), make_uint4(r[4], r[5], r[6], r[7])};
        const LeafRaw none{make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
        return big_only(leaf_decode(a, list_mode, idx < count ? raw : none));
    };
    // the keys of leaf c into P: one contiguous range, or a bucket's runs in
    // the tiles of a tile-local level (rows staged in rows_lds when staged)
    auto issue_copy = [&](const LeafDesc& c, const uint32_t* rows_lds, bool staged) {
        if (c.gather)
            gather_to_lds_async<K, NT>(P, (const K*)a.dn + c.src, a.lv.mat + c.row, c.ntiles, rows_lds, staged, tid);
        else
            copy_to_lds_async<K, NT>(P, (c.from_dn ? (const K*)a.dn : (const K*)a.d) + c.start, c.n, tid);
    };
    // Records are fetched three leaves ahead and decoded two ahead; a
    // gathered leaf's rows are staged in LDS one iteration before its keys
    // are copied, so no copy waits for a global load.
    // Leaf order.  With a counter (a.next: the pass after a tile-local level,
    // whose leaves are many and of similar size) leaves are claimed from it,
    // so a work-group that draws large leaves takes fewer: three at first,
    // then one per iteration, thread 0 claiming the leaf of iteration i + 4
    // in iteration i and publishing it in the next one, so no wave waits for
    // the atomic.  Without (leaf lists, contiguous levels: many small leaves,
    // where one counter would be contended) work-group b takes leaves
    // b, b + G, b + 2G, ... of an oversubscribed grid.
    if (count == 0) return;  // (uniform) an empty pass takes no claims
    const bool claimed = a.next != nullptr;
    const uint32_t G = gridDim.x;
    uint32_t claim = 0;
    if (tid == 0) {
        const uint32_t b = claimed ? atomicAdd(a.next, 3u) : blockIdx.x;
        const uint32_t step = claimed ? 1u : G;
        S.idx[0] = b;
        S.idx[1] = b + step;
        S.idx[2] = b + 2u * step;
        sd[0] = big_only(leaf_decode(a, list_mode, leaf_rec(a, list_mode, b, count)));
        sd[1] = big_only(leaf_decode(a, list_mode, leaf_rec(a, list_mode, b + step, count)));
        claim = claimed ? atomicAdd(a.next, 1u) : b + 3u * G;
    }
    lds_sync();
    fetch(S.idx[2], 0);
    lds_sync();
    {
        const LeafDesc c = uniform_desc(sd[0]);
        if (c.kind == kMine) issue_copy(c, S.rows[0], false);
        const LeafDesc c1 = uniform_desc(sd[1]);
        if (c1.kind == kMine && c1.gather) stage_rows_async(S.rows[1], a.lv.mat + c1.row, c1.ntiles, tid);
    }
    LeafDesc pend{0, 0, 0, 0, 0, 0, 0, 0, 0};  // sorted in A, not yet written out
    for (uint32_t i = 0;; ++i) {
        vm_wait_all();  // keys of g_i in P, rows of g_{i+1}, record of g_{i+2}, stores of g_{i-2}, the claim
        if (tid == 0) {
            S.idx[(i + 3) % 4] = claim;  // (idx[(i + 3) % 4] held g_{i-1}, read before the last barrier)
            claim = claimed ? atomicAdd(a.next, 1u) : claim + G;
            sd[(i + 2) % 4] = decode(S.idx[(i + 2) % 4], i % 2);
        }
        lds_sync();
        const uint32_t g = uni(S.idx[i % 4]);  // claims only grow: once past count, always
        {
            // rows of g_{i+2}, staged while g_i is sorted (rows[i % 2] was
            // read by the copy of g_i, issued in the previous iteration)
            const LeafDesc c2 = uniform_desc(sd[(i + 2) % 4]);
            if (c2.kind == kMine && c2.gather) stage_rows_async(S.rows[i % 2], a.lv.mat + c2.row, c2.ntiles, tid);
        }
        // the sorted g_{i-1} leaves A after the keys of g_i have left P: a
        // read of P waits for every store issued before it
        auto write_out_pending = [&]() {
            if (pend.kind) {
                K* out = dout + pend.start;
                constexpr int U = 8;
                for (uint32_t q0 = 0; q0 < pend.n; q0 += U * NT) {
                    K v[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) v[u] = A[min(q0 + u * NT + tid, CAP - 1u)];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (q0 + u * NT + tid < pend.n) out[q0 + u * NT + tid] = from_key<DT>(v[u]);
                }
                pend.kind = 0;
            }
        };
        // 6. (of g_{i-1}) buckets above the register capacity: one wave each
        //    up to bitonic_cutoff, larger ones by the work-group.  Done here,
        //    behind the barrier that opens the iteration, rather than behind
        //    a barrier of their own: s_nbig holds every add of g_{i-1}'s
        //    sort (thread 0 clears it after the next barrier).
        auto finish_pending = [&]() {
            const uint32_t nbig = s_nbig;
            if (pend.kind && nbig) {  // uniform
                const uint32_t n = pend.n;
                uint32_t nb = 2;
                while (nb < (uint32_t)LL && nb * a.sort_threshold < n) nb <<= 1;
                for (uint32_t q = wave; q < nbig; q += NW) {
                    const uint32_t b = bigs16[q];
                    const uint32_t lo = bst[b], sz = (b + 1 < nb ? bst[b + 1] : n) - lo;
                    if (sz <= a.bitonic_cutoff) {
                        K* As = A + lo;
                        wave_sort_any<K>(sz, [&](uint32_t x) { return As[x]; }, [&](uint32_t x, K v) { As[x] = v; });
                    }
                }
                lds_sync();
                for (uint32_t q = 0; q < nbig; ++q) {
                    const uint32_t b = bigs16[q];
                    const uint32_t lo = bst[b], sz = (b + 1 < nb ? bst[b + 1] : n) - lo;
                    if (sz > a.bitonic_cutoff) lds_bitonic<K>(A + lo, sz, tid, NT);  // ends with a barrier
                }
            }
        };
        if (g >= count) {
            finish_pending();
            write_out_pending();
            break;
        }
        const LeafDesc cur = uniform_desc(sd[i % 4]);
        const LeafDesc nxt = uniform_desc(sd[(i + 1) % 4]);
        // 1. keys of g_i from P into registers, with their minimum and maximum
        K kv[KPT];
        if (cur.kind == kMine) {
            K mn = key_max<K>(), mx = K(0);
            // every slot is read (P holds CAP keys), so the reads issue
            // together; slots past the leaf are masked afterwards
#pragma unroll
            for (int j = 0; j < KPT; ++j) kv[j] = P[j * NT + tid];
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                if ((uint32_t)j * NT < cur.n) {  // uniform; later steps skip these slots too
                    const bool in = j * NT + tid < cur.n;
                    const K v = cur.raw ? to_key<DT>(kv[j]) : kv[j];
                    kv[j] = in ? v : key_max<K>();
                    mn = kmin(mn, kv[j]);
                    mx = kmax(mx, in ? v : K(0));
                }
            }
            wave_minmax<K, 32>(mn, mx);
            if (lane == 0) {
                wmin[wave] = mn;
                wmax[wave] = mx;
            }
        }
        finish_pending();
        write_out_pending();
        if (tid < LL) hist[tid] = 0;
        lds_sync();  // P consumed and A written out by every wave
        if (tid == 0) s_nbig = 0;  // (read by every wave before that barrier)
        if (nxt.kind == kMine) issue_copy(nxt, S.rows[(i + 1) % 2], true);
        fetch(uni(S.idx[(i + 3) % 4]), (i + 1) % 2);  // decoded at the top of iteration i + 1
        if (cur.kind == 5) {
            // an oversize bucket of a tile-local level: its runs in order to
            // d[start, start + n), where the next level partitions it again
            gather_copy_global<K, NT>(dout + cur.start, (const K*)a.dn + cur.src, a.lv.mat + cur.row, cur.ntiles, tid);
            continue;
        }
        if (cur.kind == 3) {
            // Unbounded fallback: bitonic network in place in the source buffer.
            K* gsrc = (K*)((cur.from_dn ? (const K*)a.dn : (const K*)a.d) + cur.start);
            if (cur.raw)
                for (uint32_t q = tid; q < cur.n; q += NT) gsrc[q] = to_key<DT>(gsrc[q]);
            __threadfence_block();
            __syncthreads();
            global_bitonic<K>(gsrc, cur.n, tid, NT);
            K* out = dout + cur.start;
            for (uint32_t q = tid; q < cur.n; q += NT) out[q] = from_key<DT>(gsrc[q]);
            if (tid == 0) atomicAdd(&a.lv.ctl->oversize, 1u);
            continue;
        }
        if (cur.kind != kMine) continue;
        const uint32_t n = cur.n;
        K mn = wmin[0], mx = wmax[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) {
            mn = kmin(mn, wmin[w]);
            mx = kmax(mx, wmax[w]);
        }
        if (mn == mx) {  // constant leaf
            const K v = from_key<DT>(mn);
            K* out = dout + cur.start;
            for (uint32_t q = tid; q < n; q += NT) out[q] = v;
            continue;
        }
        // 2. local buckets by interpolation between min and max, about
        //    sort_threshold keys each.  Every key takes its rank inside its
        //    bucket from a returning LDS add (order inside a bucket is free:
        //    the bucket is sorted next) and keeps it with the bucket index as
        //    b << kRankBits | rank (b < LL, rank < CAP = 2^kRankBits), so the
        //    placement needs neither a second atomic nor the bucket again
        uint32_t nb = 2;
        while (nb < (uint32_t)LL && nb * a.sort_threshold < n) nb <<= 1;
        // b = (key - min) * nb / (range + 1) in float.  64-bit offsets are
        // first shifted down to 32 significant bits (a monotone map; u64 ->
        // double conversion has no instruction), so buckets stay monotone in
        // the key and the clamp keeps them below nb.
        const K rng = mx - mn;
        uint32_t sh = 0;
        if constexpr (sizeof(K) == 8) {
            const uint32_t rhi = (uint32_t)(rng >> 32);
            sh = rhi ? 32u - (uint32_t)__builtin_clz(rhi) : 0u;
        }
        const float scale = (float)nb / ((float)(uint32_t)(rng >> sh) + 1.0f);
        auto local_bucket = [&](K v) -> uint32_t {
            const uint32_t b = (uint32_t)((float)(uint32_t)((v - mn) >> sh) * scale);
            return min(b, nb - 1);
        };
        static_assert(KPT % 4 == 0, "rank adds go in fours");
        constexpr int kRankBits = Log2<(int)CAP>::v;
        static_assert((1u << kRankBits) == CAP && ((uint64_t)LL << kRankBits) <= (1ull << 32),
                      "bucket and rank share one word");
        uint32_t br[KPT];
#pragma unroll
        for (int j = 0; j < KPT; ++j) br[j] = 0;
#pragma unroll
        for (int j = 0; j < KPT; j += 4) {
            if ((uint32_t)j * NT < n) {  // uniform: some lane has a key here
                uint32_t ad[4], rk[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    br[j + u] = local_bucket(kv[j + u]);
                    ad[u] = (j + u) * NT + tid < n ? lds_addr(&hist[br[j + u]]) : lds_addr(&S.junk[lane]);
                }
                lds_add_rtn4(ad[0], ad[1], ad[2], ad[3], 1u, rk[0], rk[1], rk[2], rk[3]);
#pragma unroll
                for (int u = 0; u < 4; ++u) br[j + u] = (br[j + u] << kRankBits) | rk[u];
            }
        }
        lds_sync();
        // 3. bucket starts: thread t owns local bucket t
        const uint32_t mycnt = (tid < nb) ? hist[tid] : 0u;
        uint32_t tot;
        const uint32_t mystart = block_excl_scan<NT, true, false>(mycnt, scanw, &tot);  // scanw is next written two barriers on
        if (tid < nb) bst[tid] = mystart;
        lds_sync();
        // 4. place keys by bucket
#pragma unroll
        for (int j = 0; j < KPT; ++j)
            if (j * NT + tid < n) A[bst[br[j] >> kRankBits] + (br[j] & (CAP - 1u))] = kv[j];
        lds_sync();
        // 5. each thread sorts its bucket in registers with an odd-even merge
        //    network sized to the largest bucket of its wave
        {
            const bool fits = mycnt <= lane_sort_max<K>();
            uint32_t wm = fits ? mycnt : 0u;
#pragma unroll
            for (int dd = 32; dd >= 1; dd >>= 1) wm = max(wm, (uint32_t)__shfl_xor(wm, dd, 64));
            const uint32_t c = fits ? mycnt : 0u;
            // network pruned to the wave's largest bucket (rounded up to a
            // multiple of 4 above 16 keys)
            if constexpr (sizeof(K) == 4) {
                if (wm > 28) lane_sort<K, 32>(A, mystart, c);
                else if (wm > 24) lane_sort<K, 32, 28>(A, mystart, c);
                else if (wm > 20) lane_sort<K, 32, 24>(A, mystart, c);
                else if (wm > 16) lane_sort<K, 32, 20>(A, mystart, c);
                else if (wm > 12) lane_sort<K, 16>(A, mystart, c);
                else if (wm > 8) lane_sort<K, 16, 12>(A, mystart, c);
                else if (wm > 4) lane_sort<K, 8>(A, mystart, c);
                else if (wm > 1) lane_sort<K, 4>(A, mystart, c);
            } else {
                if (wm > 12) lane_sort<K, 16>(A, mystart, c);
                else if (wm > 8) lane_sort<K, 16, 12>(A, mystart, c);
                else if (wm > 4) lane_sort<K, 8>(A, mystart, c);
                else if (wm > 1) lane_sort<K, 4>(A, mystart, c);
            }
            if (!fits) bigs16[lds_add_rtn(&s_nbig, 1u)] = (uint16_t)tid;
        }
        pend = cur;  // finished (step 6) and written out in the next iteration
    }
}


}  // namespace gqs
