// oracle.cpp -- CPU oracle for the GPU-Quicksort hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load it, and only as a checker or as the timed CPU baseline.
//
// Contents and what pins each part (the tests under tests/ hold the checks):
//   * orc_std_sort_*        std::sort of a copy: the reference's own oracle
//                           (CL12/QuicksortMain.cpp:498-516, FSY :1120-1138).
//                           Pinned against numpy.sort and the committed golden
//                           fixtures (tests/golden/).
//   * orc_par_sort_*        multi-threaded CPU sort, the stand-in for
//                           tbb::parallel_sort (CL12/QuicksortMain.cpp:428-437),
//                           whose headers are absent.  Pinned against std::sort.
//   * orc_median_*          device median-of-3 (CL12/Quicksort.h:59-73) and the
//                           host median with its min/max slip
//                           (CL12/Quicksort.h:34-57).  Pinned by brute force
//                           over every ordering of 3 keys with ties, and the
//                           host one against the reference header compiled
//                           as-is into oracle/_ref.
//   * orc_optp              MAXSEQ closed form (CL12/QuicksortMain.cpp:293-295).
//                           Pinned against the table of N -> MAXSEQ values and
//                           against the closed form 2^round(log2(kN+m)).
//   * orc_f32_key/..        order-preserving float->unsigned bijection used by
//                           the GPU path.  Pinned by round trips, by fixed
//                           points (+-0, +-inf, 1.0) and by order agreement with
//                           numpy on random pairs.
//   * orc_gpuqsort_*        CPU restatement of GPUQSort: the CL12 host level
//                           loop (CL12/QuicksortMain.cpp:297-384), gqsort_kernel
//                           (CL12/QuicksortKernels.cl:102-240), lqsort_kernel
//                           (:267-421), sort_threshold/bitonic_sort (:40-97),
//                           run under one fixed legal schedule (work-groups in
//                           index order, work-items in id order).  Pinned: its
//                           output equals std::sort on every generator, and its
//                           per-level invariants (partition predicate, pivot
//                           fill, multiset conservation) are checked per level.
//   * orc_kway_classify_*   brute-force bucket classification for a k-way
//                           partition with equality buckets (linear scan over
//                           the splitters, not a binary search).  Pinned by the
//                           single-pivot case, where it reduces to the closed
//                           forms #{<p}, #{=p}, #{>p} of gqsort_kernel
//                           (CL12/QuicksortKernels.cl:135-187).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

extern "C" {

// ---------------------------------------------------------------- std::sort
void orc_std_sort_u32(uint32_t* a, size_t n) { std::sort(a, a + n); }
void orc_std_sort_f32(float* a, size_t n) { std::sort(a, a + n); }
void orc_std_sort_f64(double* a, size_t n) { std::sort(a, a + n); }
void orc_std_sort_u64(uint64_t* a, size_t n) { std::sort(a, a + n); }

// Bitwise mismatch count between two arrays of `bytes_per_key`-wide keys.
// Returns the number of differing keys; *first receives the first index
// (or n when equal).
size_t orc_mismatch(const void* a, const void* b, size_t n, size_t bytes_per_key, size_t* first) {
    const unsigned char* pa = static_cast<const unsigned char*>(a);
    const unsigned char* pb = static_cast<const unsigned char*>(b);
    size_t cnt = 0, f = n;
    for (size_t i = 0; i < n; ++i) {
        if (std::memcmp(pa + i * bytes_per_key, pb + i * bytes_per_key, bytes_per_key) != 0) {
            if (f == n) f = i;
            ++cnt;
        }
    }
    if (first) *first = f;
    return cnt;
}

}  // extern "C"

// ------------------------------------------------------- parallel CPU sort
// Sort `threads` equal chunks with std::sort, then merge pairwise in rounds.
template <class T>
static void par_sort(T* a, size_t n, int threads) {
    if (threads <= 1 || n < 4096) { std::sort(a, a + n); return; }
    std::vector<size_t> bounds(threads + 1);
    for (int t = 0; t <= threads; ++t) bounds[t] = n * (size_t)t / (size_t)threads;
    {
        std::vector<std::thread> ws;
        for (int t = 0; t < threads; ++t)
            ws.emplace_back([=, &bounds] { std::sort(a + bounds[t], a + bounds[t + 1]); });
        for (auto& w : ws) w.join();
    }
    std::vector<T> tmp(n);
    T* src = a;
    T* dst = tmp.data();
    std::vector<size_t> cur = bounds;
    while (cur.size() > 2) {
        std::vector<size_t> nxt;
        std::vector<std::thread> ws;
        for (size_t i = 0; i + 1 < cur.size(); i += 2) {
            size_t lo = cur[i];
            size_t mid = cur[i + 1];
            size_t hi = (i + 2 < cur.size()) ? cur[i + 2] : mid;
            ws.emplace_back([=] { std::merge(src + lo, src + mid, src + mid, src + hi, dst + lo); });
            nxt.push_back(lo);
            if (i + 2 >= cur.size()) break;
        }
        for (auto& w : ws) w.join();
        nxt.push_back(n);
        std::swap(src, dst);
        cur.swap(nxt);
    }
    if (src != a) std::memcpy(a, src, n * sizeof(T));
}

extern "C" {
void orc_par_sort_u32(uint32_t* a, size_t n, int threads) { par_sort(a, n, threads); }
void orc_par_sort_f32(float* a, size_t n, int threads) { par_sort(a, n, threads); }
void orc_par_sort_f64(double* a, size_t n, int threads) { par_sort(a, n, threads); }

// ------------------------------------------------------------ medians
// Device median-of-3: CL12/Quicksort.h:59-73.  select(a, b, c) returns b when
// c holds, so the two inner selects pick max(x1,x3) and max(x2,x3).
uint32_t orc_median_u32(uint32_t x1, uint32_t x2, uint32_t x3) {
    if (x1 < x2) {
        if (x2 < x3) return x2;
        return (x1 < x3) ? x3 : x1;
    }
    if (x1 < x3) return x1;
    return (x2 < x3) ? x3 : x2;
}

// Host median-of-3 as written in CL12/Quicksort.h:34-57: when x1 is the
// largest it returns min(x2, x3) instead of max(x2, x3).  Only the first
// pivot of the host level loop uses it (CL12/QuicksortMain.cpp:318).
uint32_t orc_host_median_u32(uint32_t x1, uint32_t x2, uint32_t x3) {
    if (x1 < x2) {
        if (x2 < x3) return x2;
        return (x1 < x3) ? x3 : x1;
    }
    if (x1 < x3) return x1;
    return (x2 < x3) ? x2 : x3;
}

// MAXSEQ heuristic: CL12/QuicksortMain.cpp:293-295.
size_t orc_optp(size_t s, double k, size_t m) {
    return (size_t)std::pow(2.0, std::floor(std::log((double)s * k + (double)m) / std::log(2.0) + 0.5));
}

// ------------------------------------------------------- key transforms
// Order-preserving bijection float bits -> unsigned: negative values have all
// bits flipped, non-negative values have the sign bit set.
uint32_t orc_f32_key(uint32_t u) { return (u >> 31) ? ~u : (u | 0x80000000u); }
uint32_t orc_f32_unkey(uint32_t k) { return (k >> 31) ? (k & 0x7FFFFFFFu) : ~k; }
uint64_t orc_f64_key(uint64_t u) { return (u >> 63) ? ~u : (u | 0x8000000000000000ull); }
uint64_t orc_f64_unkey(uint64_t k) { return (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k; }

}  // extern "C"

// ------------------------------------------------- GPUQSort restatement
namespace {

constexpr uint32_t kEmptyRecord = 42;  // CL12/Quicksort.h:100

template <class T>
struct WorkRec {  // CL12/Quicksort.h:105-116 (FSY/Quicksort.h:146-156 for T)
    uint32_t start = 0, end = 0;
    T pivot = T(0);
    uint32_t direction = kEmptyRecord;
};
struct ParentRec {  // CL12/Quicksort.h:124-130
    uint32_t sstart, send, oldstart, oldend, blockcount;
};
template <class T>
struct BlockRec {  // CL12/Quicksort.h:134-145
    uint32_t start, end;
    T pivot;
    uint32_t direction, parent;
};

template <class T>
T dev_median(T x1, T x2, T x3) {  // CL12/Quicksort.h:59-73 for any T
    if (x1 < x2) {
        if (x2 < x3) return x2;
        return (x1 < x3) ? x3 : x1;
    }
    if (x1 < x3) return x1;
    return (x2 < x3) ? x3 : x2;
}
template <class T>
T host_median(T x1, T x2, T x3) {  // CL12/Quicksort.h:34-57 for any T
    if (x1 < x2) {
        if (x2 < x3) return x2;
        return (x1 < x3) ? x3 : x1;
    }
    if (x1 < x3) return x1;
    return (x2 < x3) ? x2 : x3;
}

// Padding value for sort_threshold (CL12/QuicksortKernels.cl:85 pads with
// UINT_MAX).  For floating point the greatest value of the order is +inf;
// numeric_limits<T>::max() (FSY/QuicksortMain.cpp:368) would misplace +inf keys.
template <class T>
T pad_value() {
    if (std::numeric_limits<T>::has_infinity) return std::numeric_limits<T>::infinity();
    return std::numeric_limits<T>::max();
}

struct Tunables {
    uint32_t qbs, gq_wg, lq_wg, st;
};

}  // namespace

extern "C" {
// Per-run statistics of the restatement: used to size device work queues and
// to compare level counts against the GPU path.
struct orc_gpuqsort_stats {
    uint64_t levels;          // host level-loop iterations (gqsort launches)
    uint64_t blocks_total;    // sum of blocks over all levels
    uint64_t max_blocks;      // largest level
    uint64_t done_records;    // records handed to lqsort
    uint64_t max_done_len;    // largest lqsort record
    uint64_t maxseq;          // MAXSEQ used
    uint64_t invariant_violations;  // per-level partition checks that failed
};
}

namespace {

// Bitonic network over 2*WG keys, exactly CL12/QuicksortKernels.cl:40-68,
// with the work-items of each stage run in id order.
template <class T>
void ref_bitonic(T* sh, uint32_t wg) {
    for (uint32_t ulevel = 1; ulevel < wg; ulevel <<= 1) {
        for (uint32_t j = ulevel; j > 0; j >>= 1) {
            for (uint32_t lid = 0; lid < wg; ++lid) {
                uint32_t pos = 2 * lid - (lid & (j - 1));
                uint32_t direction = lid & ulevel;
                T av = sh[pos], bv = sh[pos + j];
                bool sortThem = av > bv;
                T greater = sortThem ? av : bv;
                T lesser = sortThem ? bv : av;
                sh[pos] = direction ? greater : lesser;
                sh[pos + j] = direction ? lesser : greater;
            }
        }
    }
    for (uint32_t j = wg; j > 0; j >>= 1) {
        for (uint32_t lid = 0; lid < wg; ++lid) {
            uint32_t pos = 2 * lid - (lid & (j - 1));
            T av = sh[pos], bv = sh[pos + j];
            bool sortThem = av > bv;
            sh[pos] = sortThem ? bv : av;
            sh[pos + j] = sortThem ? av : bv;
        }
    }
}

// sort_threshold: CL12/QuicksortKernels.cl:70-97.
template <class T>
void ref_sort_threshold(T* data_in, T* data_out, uint32_t start, uint32_t end, std::vector<T>& temp,
                        const Tunables& tu) {
    uint32_t tsum = end - start;
    if (tsum == tu.st) {
        ref_bitonic(data_in + start, tu.lq_wg);
        for (uint32_t i = 0; i < tu.st; ++i) data_out[start + i] = data_in[start + i];
    } else if (tsum > 1) {
        for (uint32_t i = 0; i < tu.st; ++i) temp[i] = (i < tsum) ? data_in[start + i] : pad_value<T>();
        ref_bitonic(temp.data(), tu.lq_wg);
        for (uint32_t i = 0; i < tsum; ++i) data_out[start + i] = temp[i];
    } else if (tsum == 1) {
        data_out[start] = data_in[start];
    }
}

// Exclusive work-group scan of per-item counts, returning the total, as the
// up-sweep/down-sweep at CL12/QuicksortKernels.cl:148-173 computes it.
uint32_t excl_scan(std::vector<uint32_t>& v) {
    uint32_t run = 0;
    for (auto& x : v) {
        uint32_t t = x;
        x = run;
        run += t;
    }
    return run;
}

// gqsort_kernel for one work-group: CL12/QuicksortKernels.cl:102-240.
template <class T>
void ref_gqsort_block(T* d, T* dn, const std::vector<BlockRec<T>>& blocks, std::vector<ParentRec>& parents,
                      std::vector<WorkRec<T>>& result, uint32_t blockid, const Tunables& tu,
                      uint64_t* violations) {
    const BlockRec<T> block = blocks[blockid];
    const uint32_t start = block.start, end = block.end;
    const T pivot = block.pivot;
    uint32_t direction = block.direction;
    ParentRec& pp = parents[block.parent];
    T* s = (direction == 1) ? d : dn;
    T* sn = (direction == 1) ? dn : d;
    const uint32_t wg = tu.gq_wg;
    std::vector<uint32_t> lt(wg, 0), gt(wg, 0);
    for (uint32_t lid = 0; lid < wg; ++lid) {
        for (uint32_t i = start + lid; i < end; i += wg) {
            T tmp = s[i];
            if (tmp < pivot) lt[lid]++;
            if (tmp > pivot) gt[lid]++;
        }
    }
    uint32_t ltsum = excl_scan(lt);
    uint32_t gtsum = excl_scan(gt);
    uint32_t lbeg = pp.sstart;  // atomic_add(psstart, ltsum)
    pp.sstart += ltsum;
    pp.send -= gtsum;  // atomic_sub(psend, gtsum) - gtsum
    uint32_t gbeg = pp.send;
    for (uint32_t lid = 0; lid < wg; ++lid) {
        uint32_t lfrom = lbeg + lt[lid];
        uint32_t gfrom = gbeg + gt[lid];
        for (uint32_t i = start + lid; i < end; i += wg) {
            T tmp = s[i];
            if (tmp < pivot) sn[lfrom++] = tmp;
            if (tmp > pivot) sn[gfrom++] = tmp;
        }
    }
    // atomic_dec(pblockcount) == 0 selects the finishing work-group.
    uint32_t old = pp.blockcount;
    pp.blockcount = old - 1;
    if (old != 0) return;
    const uint32_t sstart = pp.sstart, send = pp.send, oldstart = pp.oldstart, oldend = pp.oldend;
    for (uint32_t i = sstart; i < send; ++i) d[i] = pivot;
    // Level invariant: [oldstart,sstart) < pivot, [send,oldend) > pivot.
    for (uint32_t i = oldstart; i < sstart; ++i)
        if (!(sn[i] < pivot)) { ++*violations; break; }
    for (uint32_t i = send; i < oldend; ++i)
        if (!(sn[i] > pivot)) { ++*violations; break; }
    T lpivot = sn[oldstart];
    T gpivot = sn[oldend - 1];
    if (oldstart < sstart) lpivot = dev_median(lpivot, sn[(oldstart + sstart) >> 1], sn[sstart - 1]);
    if (send < oldend) gpivot = dev_median(sn[send], sn[(oldend + send) >> 1], gpivot);
    direction ^= 1;
    result[2 * blockid] = WorkRec<T>{oldstart, sstart, lpivot, direction};
    result[2 * blockid + 1] = WorkRec<T>{send, oldend, gpivot, direction};
}

// lqsort_kernel for one record: CL12/QuicksortKernels.cl:267-421.
template <class T>
void ref_lqsort_record(T* d, T* dn, const WorkRec<T>& block, const Tunables& tu) {
    struct StackRec {
        uint32_t start, end, direction;
    };
    const uint32_t wg = tu.lq_wg;
    const uint32_t d_offset = block.start;
    const uint32_t len = block.end - d_offset;
    std::vector<T> mys(len + 1), mysn(len + 1), temp(tu.st);
    std::vector<StackRec> stack;
    stack.push_back({0, len, 1});
    const T* src = (block.direction == 1) ? d : dn;
    for (uint32_t i = 0; i < len; ++i) mys[i] = src[i + d_offset];
    T* dout = d + d_offset;
    std::vector<uint32_t> lt(wg), gt(wg);
    while (!stack.empty()) {
        StackRec wr = stack.back();
        stack.pop_back();
        uint32_t start = wr.start, end = wr.end, direction = wr.direction;
        T* s = (direction == 1) ? mys.data() : mysn.data();
        T* sn = (direction == 1) ? mysn.data() : mys.data();
        T pivot = s[start];
        if (start < end) pivot = dev_median(pivot, s[(start + end) >> 1], s[end - 1]);
        std::fill(lt.begin(), lt.end(), 0u);
        std::fill(gt.begin(), gt.end(), 0u);
        for (uint32_t lid = 0; lid < wg; ++lid) {
            for (uint32_t i = start + lid; i < end; i += wg) {
                T tmp = s[i];
                if (tmp < pivot) lt[lid]++;
                if (tmp > pivot) gt[lid]++;
            }
        }
        // gfrom = end - gt[localid+1] uses the inclusive prefix (:387).
        std::vector<uint32_t> gtincl(wg);
        uint32_t acc = 0;
        for (uint32_t lid = 0; lid < wg; ++lid) { acc += gt[lid]; gtincl[lid] = acc; }
        uint32_t ltsum = excl_scan(lt);
        uint32_t gtsum = acc;
        for (uint32_t lid = 0; lid < wg; ++lid) {
            uint32_t lfrom = start + lt[lid];
            uint32_t gfrom = end - gtincl[lid];
            for (uint32_t i = start + lid; i < end; i += wg) {
                T tmp = s[i];
                if (tmp < pivot) sn[lfrom++] = tmp;
                if (tmp > pivot) sn[gfrom++] = tmp;
            }
        }
        for (uint32_t i = start + ltsum; i < end - gtsum; ++i) dout[i] = pivot;
        if (ltsum <= tu.st) ref_sort_threshold(sn, dout, start, start + ltsum, temp, tu);
        else stack.push_back({start, start + ltsum, direction ^ 1});
        if (gtsum <= tu.st) ref_sort_threshold(sn, dout, end - gtsum, end, temp, tu);
        else stack.push_back({end - gtsum, end, direction ^ 1});
    }
}

// GPUQSort host level loop: CL12/QuicksortMain.cpp:297-384 (FSY :835-906).
template <class T>
int ref_gpuqsort(T* d, T* dn, size_t size, const Tunables& tu, orc_gpuqsort_stats* st) {
    orc_gpuqsort_stats z{};
    if (!st) st = &z;
    *st = orc_gpuqsort_stats{};
    if (size == 0) return 0;
    if (size >= (size_t)UINT32_MAX) return -1;  // uint32 indexing (CL12/QuicksortMain.cpp:405)
    if (tu.st != 2 * tu.lq_wg) return -2;        // CL12/Quicksort.h:77
    const size_t MAXSEQ = orc_optp(size, 0.00009516, 203);
    st->maxseq = MAXSEQ;
    T startpivot = host_median(d[0], d[size / 2], d[size - 1]);
    std::vector<WorkRec<T>> work, done, news;
    std::vector<ParentRec> parents;
    std::vector<BlockRec<T>> blocks;
    work.push_back(WorkRec<T>{0, (uint32_t)size, startpivot, 1});
    while (!work.empty()) {
        size_t blocksize = 0;
        for (auto& w : work) blocksize += std::max<size_t>((w.end - w.start) / MAXSEQ, 1);
        for (auto& w : work) {
            uint32_t blockcount = (uint32_t)((w.end - w.start + blocksize - 1) / blocksize);
            parents.push_back(ParentRec{w.start, w.end, w.start, w.end, blockcount - 1});
            for (uint32_t i = 0; i + 1 < blockcount; ++i) {
                uint32_t bstart = w.start + (uint32_t)blocksize * i;
                blocks.push_back(BlockRec<T>{bstart, bstart + (uint32_t)blocksize, w.pivot, w.direction,
                                             (uint32_t)parents.size() - 1});
            }
            blocks.push_back(BlockRec<T>{w.start + (uint32_t)blocksize * (blockcount - 1), w.end, w.pivot,
                                         w.direction, (uint32_t)parents.size() - 1});
        }
        news.assign(blocks.size() * 2, WorkRec<T>{});
        for (uint32_t b = 0; b < blocks.size(); ++b)
            ref_gqsort_block(d, dn, blocks, parents, news, b, tu, &st->invariant_violations);
        st->levels++;
        st->blocks_total += blocks.size();
        st->max_blocks = std::max<uint64_t>(st->max_blocks, blocks.size());
        work.clear();
        parents.clear();
        blocks.clear();
        for (auto& r : news) {
            if (r.direction == kEmptyRecord) continue;
            if (r.end - r.start <= tu.qbs) {
                if (r.end - r.start > 0) done.push_back(r);
            } else {
                work.push_back(r);
            }
        }
        news.clear();
    }
    for (auto& r : done) {
        st->max_done_len = std::max<uint64_t>(st->max_done_len, r.end - r.start);
        ref_lqsort_record(d, dn, r, tu);
    }
    st->done_records = done.size();
    return 0;
}

}  // namespace

extern "C" {
// CPU restatement of GPUQSort.  d: in/out, dn: scratch of n keys.  Tunables as
// in CL12/Quicksort.h:77-98 (Intel GPU set: 1728, 128, 128, 256).
int orc_gpuqsort_u32(uint32_t* d, uint32_t* dn, size_t n, uint32_t qbs, uint32_t gq_wg, uint32_t lq_wg,
                     uint32_t st, orc_gpuqsort_stats* stats) {
    return ref_gpuqsort<uint32_t>(d, dn, n, Tunables{qbs, gq_wg, lq_wg, st}, stats);
}
int orc_gpuqsort_f32(float* d, float* dn, size_t n, uint32_t qbs, uint32_t gq_wg, uint32_t lq_wg, uint32_t st,
                     orc_gpuqsort_stats* stats) {
    return ref_gpuqsort<float>(d, dn, n, Tunables{qbs, gq_wg, lq_wg, st}, stats);
}
int orc_gpuqsort_f64(double* d, double* dn, size_t n, uint32_t qbs, uint32_t gq_wg, uint32_t lq_wg, uint32_t st,
                     orc_gpuqsort_stats* stats) {
    return ref_gpuqsort<double>(d, dn, n, Tunables{qbs, gq_wg, lq_wg, st}, stats);
}

}  // extern "C"

// ------------------------------------------ arrays of ordered keys / segmented std::sort
// orc_keys_*: the float -> ordered-key bijection (orc_f32_key / orc_f64_key)
// over an array, for comparing the GPU's level output with its input.
// orc_std_sort_segmented_*: std::sort of each array [off[i], off[i+1]) on
// its own, the per-array oracle of a segmented sort.
template <class T>
static void std_sort_segmented(T* a, const uint64_t* off, size_t nseg) {
    for (size_t i = 0; i < nseg; ++i) std::sort(a + off[i], a + off[i + 1]);
}

extern "C" {
void orc_keys_f32(const uint32_t* in, uint32_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = orc_f32_key(in[i]);
}
void orc_keys_f64(const uint64_t* in, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = orc_f64_key(in[i]);
}
void orc_std_sort_segmented_u32(uint32_t* a, const uint64_t* off, size_t nseg) { std_sort_segmented(a, off, nseg); }
void orc_std_sort_segmented_f32(float* a, const uint64_t* off, size_t nseg) { std_sort_segmented(a, off, nseg); }
void orc_std_sort_segmented_f64(double* a, const uint64_t* off, size_t nseg) { std_sort_segmented(a, off, nseg); }
}  // extern "C"

// ------------------------------------------ k-way classification (brute force)
// Splitters u[0..m-1] strictly increasing; eq[i] != 0 gives u[i] its own
// equality bucket.  Bucket order: (-inf,u0) [=u0] (u0,u1) [=u1] ... (u_{m-1},+inf),
// where a range bucket (u_{i-1}, u_i] also holds keys == u_i when eq[i] == 0.
// Writes each key's bucket id and adds to counts[nbuckets].  Returns nbuckets.
template <class K>
static int kway_classify(const K* keys, size_t n, const K* u, const uint8_t* eq, int m, uint16_t* bucket_out,
                         uint64_t* counts) {
    std::vector<int> id_range(m + 1), id_eq(m, -1);
    int next = 0;
    for (int i = 0; i <= m; ++i) {
        id_range[i] = next++;
        if (i < m && eq[i]) id_eq[i] = next++;
    }
    for (size_t j = 0; j < n; ++j) {
        K k = keys[j];
        int b = -1;
        for (int i = 0; i < m && b < 0; ++i) {
            if (k < u[i]) b = id_range[i];
            else if (k == u[i]) b = eq[i] ? id_eq[i] : id_range[i];
        }
        if (b < 0) b = id_range[m];
        if (bucket_out) bucket_out[j] = (uint16_t)b;
        if (counts) counts[b]++;
    }
    return next;
}

extern "C" {
int orc_kway_classify_u32(const uint32_t* keys, size_t n, const uint32_t* u, const uint8_t* eq, int m,
                          uint16_t* bucket_out, uint64_t* counts) {
    return kway_classify(keys, n, u, eq, m, bucket_out, counts);
}
int orc_kway_classify_u64(const uint64_t* keys, size_t n, const uint64_t* u, const uint8_t* eq, int m,
                          uint16_t* bucket_out, uint64_t* counts) {
    return kway_classify(keys, n, u, eq, m, bucket_out, counts);
}

}  // extern "C"
