// merge.hip -- pairwise merge of sorted runs, 64-bit indices throughout.
//
// A sort of n >= 2^32 keys (beyond the 32-bit segment sizes the partition
// levels use) sorts runs of at most 2^31 keys with the level machinery and
// then merges the runs pairwise, one pass over the keys per round.  A merge
// follows Merge Path (O. Green, R. McColl, D. A. Bader, "GPU Merge Path: A GPU
// Merging Algorithm", ICS 2012): output position p of merge(A, B) holds i(p)
// keys of A and p - i(p) keys of B, where i(p) is the smallest i with
// A[i] > B[p-1-i] (A wins ties), found by binary search along the cross
// diagonal through p.
//   k_merge_split  i(p) at every tile boundary p = t * tile (one thread each);
//   k_merge_tile   one work-group per tile of outputs: it stages
//                  A[i(p0), i(p1)) and B[p0-i(p0), p1-i(p1)) in LDS as ordered
//                  keys, each thread finds its own diagonal inside the tile
//                  the same way and merges its KPT outputs sequentially, and
//                  the tile goes back to HBM through LDS as coalesced stores.
// Float keys are compared through the same order-preserving key transform as
// the sort, so merging sorted runs gives the sort's order.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "keys.h"

namespace gqs {

constexpr int kMergeWG = 256;
template <class K>
constexpr int merge_kpt() { return sizeof(K) == 4 ? 16 : 8; }
template <class K>
constexpr uint32_t merge_tile() { return kMergeWG * merge_kpt<K>(); }

template <int DT>
__global__ __launch_bounds__(kMergeWG) void k_merge_split(const typename KeyOf<DT>::K* A, uint64_t na,
                                                         const typename KeyOf<DT>::K* B, uint64_t nb,
                                                         uint64_t* split, uint64_t nsplit) {
    using K = typename KeyOf<DT>::K;
    const uint64_t t = (uint64_t)blockIdx.x * kMergeWG + threadIdx.x;
    if (t >= nsplit) return;
    const uint64_t p = min(t * (uint64_t)merge_tile<K>(), na + nb);
    uint64_t lo = p > nb ? p - nb : 0, hi = min(p, na);
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;  // lo <= mid < min(p, na), so p-1-mid < nb
        if (to_key<DT>(A[mid]) <= to_key<DT>(B[p - 1 - mid])) lo = mid + 1;
        else hi = mid;
    }
    split[t] = lo;
}

template <int DT>
__global__ __launch_bounds__(kMergeWG) void k_merge_tile(const typename KeyOf<DT>::K* A, uint64_t na,
                                                        const typename KeyOf<DT>::K* B, uint64_t nb,
                                                        typename KeyOf<DT>::K* out, const uint64_t* split,
                                                        uint64_t tile0) {
    using K = typename KeyOf<DT>::K;
    constexpr int KPT = merge_kpt<K>();
    constexpr uint32_t T = merge_tile<K>();
    __shared__ K s[T];
    const uint32_t tid = threadIdx.x;
    const uint64_t g = tile0 + blockIdx.x;
    const uint64_t p0 = g * T, p1 = min(p0 + T, na + nb);
    // For sorted runs split[] is non-decreasing and the tile takes
    // split[g+1] - split[g] <= T keys of A.  The clamp changes nothing then;
    // for unsorted input (a run left unsorted by a reported overflow) it keeps
    // every index within A, B and the tile.
    const uint64_t a0 = split[g];
    const uint64_t a1 = min(max(split[g + 1], a0), a0 + (p1 - p0));
    const uint64_t b0 = p0 - a0, b1 = p1 - a1;
    const uint32_t la = (uint32_t)(a1 - a0), lb = (uint32_t)(b1 - b0), len = la + lb;  // len <= T
    for (uint32_t q = tid; q < la; q += kMergeWG) s[q] = to_key<DT>(A[a0 + q]);
    for (uint32_t q = tid; q < lb; q += kMergeWG) s[la + q] = to_key<DT>(B[b0 + q]);
    __syncthreads();
    const K* sa = s;
    const K* sb = s + la;
    // this thread's outputs are tile positions [pp, pp + KPT)
    const uint32_t pp = tid * (uint32_t)KPT;
    const uint32_t p = min(pp, len);
    uint32_t lo = p > lb ? p - lb : 0, hi = min(p, la);
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sa[mid] <= sb[p - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    uint32_t i = lo, j = p - lo;
    K v[KPT];
#pragma unroll
    for (int e = 0; e < KPT; ++e) {
        const bool a_ok = i < la, b_ok = j < lb;
        const K ka = a_ok ? sa[i] : K(0), kb = b_ok ? sb[j] : K(0);
        const bool take_a = a_ok && (!b_ok || ka <= kb);
        v[e] = take_a ? ka : kb;  // positions past len are never stored
        i += take_a ? 1u : 0u;
        j += take_a ? 0u : 1u;
    }
    __syncthreads();  // every thread has finished reading the staged runs
#pragma unroll
    for (int e = 0; e < KPT; ++e)
        if (pp + e < len) s[pp + e] = v[e];
    __syncthreads();
    for (uint32_t q = tid; q < len; q += kMergeWG) out[p0 + q] = from_key<DT>(s[q]);
}

uint64_t merge_tile_keys(int dtype) { return dtype == DT_F64 ? merge_tile<uint64_t>() : merge_tile<uint32_t>(); }

uint64_t merge_split_entries(int dtype, uint64_t n) {
    const uint64_t T = merge_tile_keys(dtype);
    return (n + T - 1) / T + 1;
}

template <int DT>
static hipError_t merge_split_t(const void* a, uint64_t na, const void* b, uint64_t nb, uint64_t* split,
                                hipStream_t st) {
    using K = typename KeyOf<DT>::K;
    const uint64_t ntiles = (na + nb + merge_tile<K>() - 1) / merge_tile<K>();
    const uint64_t nsplit = ntiles + 1;
    if (ntiles == 0) return hipSuccess;
    if (ntiles > 0x7FFFFFFFull) return hipErrorInvalidValue;
    k_merge_split<DT><<<dim3((uint32_t)((nsplit + kMergeWG - 1) / kMergeWG)), dim3(kMergeWG), 0, st>>>(
        (const K*)a, na, (const K*)b, nb, split, nsplit);
    return hipGetLastError();
}

template <int DT>
static hipError_t merge_tiles_t(const void* a, uint64_t na, const void* b, uint64_t nb, void* out,
                                const uint64_t* split, uint64_t t0, uint64_t t1, hipStream_t st) {
    using K = typename KeyOf<DT>::K;
    const uint64_t ntiles = (na + nb + merge_tile<K>() - 1) / merge_tile<K>();
    t1 = t1 < ntiles ? t1 : ntiles;
    if (t0 >= t1) return hipSuccess;
    if (t1 - t0 > 0x7FFFFFFFull) return hipErrorInvalidValue;
    k_merge_tile<DT><<<dim3((uint32_t)(t1 - t0)), dim3(kMergeWG), 0, st>>>((const K*)a, na, (const K*)b, nb,
                                                                         (K*)out, split, t0);
    return hipGetLastError();
}

hipError_t launch_merge_split(int dtype, const void* a, uint64_t na, const void* b, uint64_t nb, uint64_t* split,
                              hipStream_t st) {
    switch (dtype) {
        case DT_U32: return merge_split_t<DT_U32>(a, na, b, nb, split, st);
        case DT_F32: return merge_split_t<DT_F32>(a, na, b, nb, split, st);
        case DT_F64: return merge_split_t<DT_F64>(a, na, b, nb, split, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_merge_tiles(int dtype, const void* a, uint64_t na, const void* b, uint64_t nb, void* out,
                              const uint64_t* split, uint64_t tile_begin, uint64_t tile_end, hipStream_t st) {
    switch (dtype) {
        case DT_U32: return merge_tiles_t<DT_U32>(a, na, b, nb, out, split, tile_begin, tile_end, st);
        case DT_F32: return merge_tiles_t<DT_F32>(a, na, b, nb, out, split, tile_begin, tile_end, st);
        case DT_F64: return merge_tiles_t<DT_F64>(a, na, b, nb, out, split, tile_begin, tile_end, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_merge(int dtype, const void* a, uint64_t na, const void* b, uint64_t nb, void* out, uint64_t* split,
                        hipStream_t st) {
    hipError_t e = launch_merge_split(dtype, a, na, b, nb, split, st);
    return e == hipSuccess ? launch_merge_tiles(dtype, a, na, b, nb, out, split, 0, ~0ull, st) : e;
}

}  // namespace gqs
