/* gqsort.h -- C ABI of the MI355X GPU-Quicksort library (libgqsort.so).
 *
 * Drop-in for the reference's sort entry point
 *     template <class T> void GPUQSort(OCLResources*, size_t size, T* d, T* dn)
 *     (GPU-Quicksort_OpenCL_1.2/QuicksortMain.cpp:297-298,
 *      GPU-Quicksort_SYCL/QuicksortMain.cpp:357-358,
 *      GPU-Quicksort_Full_SYCL/QuicksortMain.cpp:835-836)
 * and the OpenCL 2.0 device-scheduled variant
 *     GPUQSort(OCLResources*, size_t size, T* d, work_record* pdone,
 *              work_record* pnews, size_t MAX_SIZE)
 *     (GPU-Quicksort_OpenCL_2.0/QuicksortMain.cpp:301-302).
 *
 * Contract kept from the reference: `d` holds the keys and receives them
 * sorted ascending, `dn` is caller-owned scratch of at least n keys that is
 * clobbered, the library never frees caller memory.  Differences: every call
 * returns a status instead of asserting (CL12/OpenCLUtils.h:41-45), and keys
 * may be uint32, float or double (the reference's OpenCL variants sort uint
 * only; the full-SYCL variant instantiates uint/float/double,
 * GPU-Quicksort_Full_SYCL/QuicksortMain.cpp:1192-1194).  Float keys are
 * ordered as IEEE values for inputs without NaN; -0.0 sorts before +0.0.
 *
 * Threading: a context serialises its own calls with an internal lock (a
 * host-pointer gqs_sort holds it from upload to download); use one context
 * per host thread for concurrency.  Streams are passed as void* (a
 * hipStream_t; NULL is the default stream).  Stream-ordered calls on one
 * context share its workspace: a call enqueued on another stream than the
 * previous call first waits (hipStreamWaitEvent) for the previous call's work,
 * so sorts on different streams of one context run one after the other.
 *
 * Status: the reference's failures are silent asserts (CL12/OpenCLUtils.h:
 * 41-45).  Here a bound hit on the device (a queue capacity, or bad segment
 * offsets) skips the affected keys, sets a sticky flag and is reported as
 * GQS_EOVERFLOW (capacity) or GQS_EINVAL (offsets): synchronously by gqs_sort,
 * and for the stream-ordered calls by the next gqs_get_stats on the context
 * (gqs_graph_status for a captured graph), which waits for the call.
 */
#ifndef GQSORT_H
#define GQSORT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gqs_status {
    GQS_OK = 0,
    GQS_EINVAL = 1,      /* null pointer, bad dtype, n out of range, bad parameter */
    GQS_ENOMEM = 2,      /* device or pinned-host allocation failed */
    GQS_EDEVICE = 3,     /* a HIP call failed (see gqs_last_error) */
    GQS_EOVERFLOW = 4,   /* an internal queue bound was exceeded; the affected keys are left unsorted */
    GQS_EBUSY = 5,       /* serving: memory budget or slots exhausted, request not admitted */
    GQS_ENOTFOUND = 6    /* serving: unknown ticket */
} gqs_status;

typedef enum gqs_dtype { GQS_U32 = 0, GQS_F32 = 1, GQS_F64 = 2 } gqs_dtype;

typedef enum gqs_mode {
    GQS_MODE_MULTIPIVOT = 0, /* sample-chosen splitters, up to 255 per segment (default) */
    GQS_MODE_CLASSIC = 1     /* one median-of-3 pivot per segment, QBS leaf cutoff (the reference's schedule) */
} gqs_mode;

typedef struct gqs_ctx gqs_ctx;

/* Tunables (the reference's compile-time QUICKSORT_BLOCK_SIZE,
 * SORT_THRESHOLD and work-group sizes, CL12/Quicksort.h:77-98, become
 * run-time parameters chosen by the autotuner). */
typedef struct gqs_params {
    int32_t mode;          /* gqs_mode */
    uint32_t leaf_max;     /* largest bucket sorted in LDS (QUICKSORT_BLOCK_SIZE role); 16..16384 (u32/f32); f64 uses at
                              most 8192.  Leaves above 8192 (f64: 4096) keys are sorted by a 1024-thread
                              work-group, smaller ones by a 512-thread one */
    uint32_t leaf_target;  /* multipivot: bucket size the fan-out aims for */
    uint32_t oversample;   /* multipivot: sample keys per bucket */
    uint32_t max_fanout;   /* multipivot: largest fan-out, power of two in [2, 256] */
    uint32_t max_levels;   /* partition levels before every bucket goes to the local sort */
    uint32_t host_driven;  /* 1: read the level queue back after each level and stop early (CL12 loop) */
    uint32_t part_wgs_per_cu; /* persistent count/scatter work-groups per CU */
    uint32_t leaf_wgs_per_cu; /* leaf-kernel work-groups per CU (after a tile-local level: two per CU) */
    uint32_t sort_threshold;  /* local sort: bucket size its partition aims for (SORT_THRESHOLD role), 4..4096 */
    uint32_t bitonic_cutoff;  /* local sort: largest sub-bucket sorted by one wave's bitonic network, 64..512 */
} gqs_params;

/* Counters of the last sort on a context. */
typedef struct gqs_stats {
    uint32_t levels_launched;   /* partition levels enqueued */
    uint32_t levels_used;       /* levels that had segments */
    uint32_t oversize_leaves;   /* leaves sorted by the unbounded fallback */
    uint32_t overflow;          /* non-zero if a queue bound was hit */
    uint32_t segments[8];       /* segments per level (first 8 levels) */
    uint64_t level_keys[8];     /* keys partitioned per level (first 8 levels) */
} gqs_stats;

/* Context: owns the device, its workspace (grown on demand) and parameters. */
gqs_status gqs_create(gqs_ctx** ctx, int device);
gqs_status gqs_destroy(gqs_ctx* ctx);
const char* gqs_last_error(const gqs_ctx* ctx);
gqs_status gqs_get_params(const gqs_ctx* ctx, gqs_params* out);
gqs_status gqs_set_params(gqs_ctx* ctx, const gqs_params* p);
gqs_status gqs_default_params(int dtype, gqs_params* out);
/* GQS_OK when every field is in range (the checks of gqs_set_params and of
 * every entry gqs_tune_load reads); needs no device. */
gqs_status gqs_check_params(const gqs_params* p);

/* Reference semantics: d, dn are HOST pointers; synchronous.  The keys are
 * copied to device buffers owned by the context, sorted and copied back.
 * From 2^24 keys (gqs_set_host_pipeline) the copies overlap the sort: the
 * input goes up in eight runs, each sorted while the next ones are in
 * flight, the runs are merged as soon as pairs are sorted, and the final
 * merge's output goes back in eight ranges, each copied while the next is
 * merged.  Pinned host arrays (gqs_host_alloc) are copied without staging.
 * On an error status d's contents are unspecified. */
gqs_status gqs_sort(gqs_ctx* ctx, int dtype, void* d, void* dn, size_t n);

/* Smallest n for which gqs_sort overlaps the copies with the sort:
 * min_keys >= 2^16, 0 for the default (2^24), SIZE_MAX to never overlap.  No
 * counterpart in the reference (its GPUQSort wraps the host arrays as device
 * buffers, sorts, then releases them, CL12/QuicksortMain.cpp:300-303,
 * 381-383); tests use it to run both host paths at sizes a check can afford. */
gqs_status gqs_set_host_pipeline(gqs_ctx* ctx, size_t min_keys);

/* Device-resident: d, dn are device pointers; stream-ordered, returns after
 * enqueueing (no host synchronisation unless params.host_driven).
 *
 * n is a size_t on both entry points.  The reference caps n at 32 bits
 * (CL12/QuicksortMain.cpp:405; uint indices in its records,
 * CL12/Quicksort.h:105-145); here n >= 2^32 - 2^16 is sorted as runs of 2^31 keys
 * (each through the partition levels) merged pairwise with 64-bit indices,
 * ping-ponging between d and dn.  gqs_get_stats then describes the last run. */
gqs_status gqs_sort_device(gqs_ctx* ctx, int dtype, void* d, void* dn, size_t n, void* stream);

/* One partition level of the keys d[0, n) (host-synchronous), without the
 * local sort: the reference's gqsort_kernel launch and the child records it
 * returns (CL12/QuicksortMain.cpp:332-373, CL12/QuicksortKernels.cl:102-240),
 * generalised to up to 255 splitters.  Keys are written to dn grouped by
 * bucket, in ordered-key representation (uint32 unchanged; float and double
 * bits mapped by the order-preserving bijection: non-negative values get the
 * sign bit set, negative values all bits flipped).  Outputs, host memory:
 * splitters[0, *nsplit) (strictly increasing, same representation, up to 255
 * keys), eq[i] = 1 when splitter i has an equality bucket, and per bucket
 * (up to 256) its first key, its size and whether it is an equality bucket.
 * Bucket order: keys up to splitter 0, [keys equal to splitter 0], keys up to
 * splitter 1, ...; an equality bucket's keys are not stored (the sort fills
 * their value in afterwards, the reference's pivot fill, :214-216). */
gqs_status gqs_partition_level(gqs_ctx* ctx, int dtype, void* d, void* dn, size_t n, void* splitters, uint8_t* eq,
                               uint32_t* nsplit, uint64_t* bucket_start, uint32_t* bucket_size, uint8_t* bucket_eq,
                               uint32_t* nbuckets);

/* No counterpart in the reference: gqs_sort_device without any allocation,
 * for callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The caller passes d
 * and dn as for gqs_sort_device; the sort's workspace comes from a 24 MiB
 * device array that each GPU the library runs on holds until the process
 * ends.  Limits: n <= 2^24 keys (GQS_EINVAL above, or when the current
 * parameters need a larger workspace); no per-launch profiling; the large-n
 * run length (gqs_set_large_run) does not apply.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (from any context, stream or
 * thread) race, so the caller orders them (one stream, or waits between). */
gqs_status gqs_sort_device_noalloc(gqs_ctx* ctx, int dtype, void* d, void* dn, size_t n, void* stream);

/* Run length of the large-n path: with keys > 0, every sort of more than
 * `keys` keys on this context takes that path with runs of `keys` keys
 * (2 <= keys < 2^32 - 2^16); keys == 0 restores the default (runs of 2^31
 * keys, only for n >= 2^32 - 2^16).  No counterpart in the reference: it lets tests
 * and tuning exercise the run-and-merge path at sizes a check can afford. */
gqs_status gqs_set_large_run(gqs_ctx* ctx, size_t keys);

/* Segmented: nseg independent arrays packed in d, array i occupying
 * [offsets[i], offsets[i+1]) (offsets: nseg+1 device uint64).  Each array is
 * sorted on its own; one chain of launches for all of them.  total_keys (and a
 * graph's max_total_keys) must be below 2^32 - 2^16. */
gqs_status gqs_sort_segmented(gqs_ctx* ctx, int dtype, void* d, void* dn, const uint64_t* offsets, size_t nseg,
                              size_t total_keys, void* stream);

/* The per-array baseline of a segmented sort: one plain gqs_sort_device call
 * per array (host_offsets: nseg+1 HOST uint64), all enqueued on `stream`. */
gqs_status gqs_sort_each(gqs_ctx* ctx, int dtype, void* d, void* dn, const uint64_t* host_offsets, size_t nseg,
                         void* stream);

/* Segmented sort captured once into a hipGraph and replayed: the graph owns
 * its workspace (sized for max_nseg arrays and max_total_keys keys), reads the
 * current number of arrays from the device word *nseg_dev and the bounds from
 * offsets[0..nseg] at every replay, and sorts the arrays of d in place (dn is
 * scratch).  d, dn, offsets and nseg_dev must stay valid while the graph lives;
 * refill them (stream-ordered) before each gqs_graph_launch. */
typedef struct gqs_graph gqs_graph;
gqs_status gqs_graph_create_segmented(gqs_ctx* ctx, int dtype, void* d, void* dn, const uint64_t* offsets,
                                      const uint32_t* nseg_dev, size_t max_nseg, size_t max_total_keys,
                                      gqs_graph** out);
gqs_status gqs_graph_launch(gqs_graph* g, void* stream);
/* Waits for the last replay and returns its status: GQS_EINVAL if an array's
 * offsets decreased or passed max_total_keys, or *nseg_dev exceeded max_nseg
 * (those arrays were skipped), GQS_EOVERFLOW on a queue bound, else GQS_OK. */
gqs_status gqs_graph_status(gqs_graph* g);
gqs_status gqs_graph_destroy(gqs_graph* g);

/* Autotuner (the reference hand-picks QUICKSORT_BLOCK_SIZE, SORT_THRESHOLD and
 * work-group sizes per vendor, CL12/Quicksort.h:77-98).  Times sorts of n
 * synthetic keys (generator `gen`, see gqs_generate) on the context's device by
 * coordinate descent over the parameters, keeps the fastest set for
 * (dtype, floor(log2 n)) in the context's table and returns it.  Sorts use the
 * table entry of their (dtype, size class), or of the nearest tuned size class
 * of the same dtype, until the next gqs_set_params.  The table is saved/loaded
 * as JSON. */
gqs_status gqs_autotune(gqs_ctx* ctx, int dtype, size_t n, int gen, uint32_t reps, gqs_params* best);
gqs_status gqs_tune_save(gqs_ctx* ctx, const char* path);
gqs_status gqs_tune_load(gqs_ctx* ctx, const char* path);

/* Statistics of the last sort (waits for it).  Returns the last call's status:
 * GQS_EOVERFLOW if a queue bound was hit, GQS_EINVAL if segment offsets were
 * out of range (out is filled either way). */
gqs_status gqs_get_stats(gqs_ctx* ctx, gqs_stats* out);

/* Testing hook, no counterpart in the reference: caps the segment queue of
 * every following sort on this context at `segments` entries (0 removes the
 * cap), so that the overflow path can be driven from the host. */
gqs_status gqs_set_queue_limit(gqs_ctx* ctx, uint32_t segments);

/* Testing hook, no counterpart in the reference: the level schedule of later
 * sorts of one range that plan two partition levels and a spare.  0 (the
 * default): both levels tile-local, the first in place, the second reading
 * the first one's runs (no count pass).  1: a contiguous level (count, scan,
 * scatter) before the tile-local one.  2: as 0, with the second level reading
 * its runs through 64-entry run tables instead of one entry per thread, so
 * that tests reach the path taken by runs too short for one table.  All give
 * the same output; the hook lets tests and benchmarks compare them. */
gqs_status gqs_set_level_scheme(gqs_ctx* ctx, int scheme);

/* Per-kernel timing with HIP events on the sort's own stream.  While enabled,
 * every launch is bracketed by events; gqs_kernel_times() synchronises and
 * returns, per kernel kind, the number of launches and total milliseconds.
 * Kinds: 0 splitters, 1 tileplan, 2 count, 3 scan, 4 plan, 5 scatter, 6 leaf, 7 init, 8 fill
 * (kept for the layout of the arrays: equality buckets are filled inside the
 * leaf launches, so kind 8 reports no launches), 9 tilepart (the tile-local
 * partition of the last planned level). */
gqs_status gqs_set_profiling(gqs_ctx* ctx, int on);
gqs_status gqs_kernel_times(gqs_ctx* ctx, uint32_t* launches, double* ms, int nkinds);
/* Per-launch (kind, ms) records collected by the last gqs_kernel_times();
 * copies up to cap records and returns how many there are (-1 on bad args). */
int gqs_kernel_trace(gqs_ctx* ctx, int* kinds, float* ms, int cap);

/* Device properties used for launch shapes (CUs, LDS per work-group, name). */
gqs_status gqs_device_info(gqs_ctx* ctx, int* cus, int* lds_per_wg, int* wave_size, char* name, size_t name_len);

/* The device report of the reference's show_CL option
 * (QueryPrintOpenCLDeviceInfo, CL12/OpenCLUtils.h:291-387: vendor, name,
 * driver version, compute units, work-item sizes, work-group size, clock,
 * local and global memory), from hipGetDeviceProperties. */
typedef struct gqs_device_props {
    char name[256];             /* product name (the ISA name when the driver reports none) */
    char arch[256];             /* ISA and target features, e.g. gfx950:sramecc+:xnack- */
    char pci_bus_id[32];        /* domain:bus:device.function */
    int32_t driver_version;     /* hipDriverGetVersion */
    int32_t runtime_version;    /* hipRuntimeGetVersion */
    int32_t compute_units;
    int32_t wave_size;
    int32_t max_threads_per_block;
    int32_t max_block_dim[3];
    int32_t clock_khz;          /* peak shader clock */
    int32_t mem_clock_khz;      /* peak memory clock */
    int32_t mem_bus_bits;       /* memory interface width */
    int32_t reserved;
    uint64_t lds_per_block;     /* bytes of LDS one work-group may use */
    uint64_t lds_per_cu;        /* bytes of LDS per compute unit */
    uint64_t l2_bytes;
    uint64_t global_mem_bytes;  /* HBM capacity */
} gqs_device_props;
gqs_status gqs_device_report(gqs_ctx* ctx, gqs_device_props* out);

/* Synthetic inputs for the benchmark configurations, generated on the host
 * with a counter-based hash (deterministic for a seed, independent of thread
 * count).  gen: 0 uniform, 1 gaussian (mean of 4 uniforms), 2 sorted,
 * 3 reverse sorted, 4 all equal, 5 bucket (256 shuffled runs), 6 permutation of
 * 1..n (the reference's own data, CL12/QuicksortMain.cpp:415-416),
 * 7 normal floats with 1% wide outliers and 0.1% duplicates. */
gqs_status gqs_generate(int dtype, int gen, uint64_t seed, void* out, size_t n);

/* ---- Serving: continuous batching of sort requests ------------------------
 * A server owns a context, `slots` batch slots each with device-resident key
 * buffers of slot_keys 8-byte keys and a sort workspace (all inside the HBM
 * budget), one upload, one sort and one download stream, and a scheduler
 * thread.  Whenever a slot is free the scheduler packs the oldest
 * pending requests of one dtype (about batch_bytes, up to the slot's
 * capacity) into one segmented sort: upload, sort, download, with the
 * upload of one batch, the sort of the previous one and the download of the
 * one before running at once.  A request is done when its own keys are back.
 * in/out are host buffers of n keys (pinned memory gives overlapped copies;
 * in == out allowed); they must stay valid until the request is done.
 * Destroy a server only when no thread waits on it. */
typedef struct gqs_server gqs_server;
typedef struct gqs_server_config {
    uint64_t hbm_budget_bytes;   /* device memory the server may hold (default 64 GiB) */
    uint64_t slot_keys;          /* keys (8-byte units) per slot buffer; 0: budget / (24 * slots) */
    uint32_t slots;              /* batches in flight, 1..16 (default 8) */
    uint32_t max_batch_requests; /* requests per batch (default 4096) */
    uint64_t batch_bytes;        /* batch target: requests join the oldest one's batch while the
                                    batch stays within it (0: 64 MiB) */
} gqs_server_config;
typedef struct gqs_server_stats {
    uint64_t submitted, completed, failed, batches, max_batch_requests, keys_sorted, pending;
} gqs_server_stats;
gqs_status gqs_server_default_config(gqs_server_config* cfg);
/* The configuration a server created from `cfg` runs with (defaults filled,
 * slot_keys derived from the budget when 0 and capped so that a batch of
 * 4-byte keys stays below 2^32 - 2^16 keys); needs no device. */
gqs_status gqs_server_resolve_config(const gqs_server_config* cfg, gqs_server_config* out);
gqs_status gqs_server_create(int device, const gqs_server_config* cfg, gqs_server** out);
/* GQS_EBUSY: the request can never fit a slot. */
gqs_status gqs_server_submit(gqs_server* s, int dtype, const void* in, void* out, size_t n, uint64_t* ticket);
/* timeout_us < 0 waits forever; GQS_EBUSY on timeout, GQS_EDEVICE if the request failed. */
gqs_status gqs_server_wait(gqs_server* s, uint64_t ticket, int64_t timeout_us);
gqs_status gqs_server_poll(gqs_server* s, uint64_t ticket, int* done);
/* Steady-clock microseconds of submission, admission into a batch, completion. */
gqs_status gqs_server_request_times(gqs_server* s, uint64_t ticket, double* submit_us, double* admit_us,
                                    double* done_us);
gqs_status gqs_server_release(gqs_server* s, uint64_t ticket); /* forget a finished request */
gqs_status gqs_server_get_stats(gqs_server* s, gqs_server_stats* out);
gqs_status gqs_server_get_config(const gqs_server* s, gqs_server_config* out); /* effective (derived) values */
gqs_status gqs_server_destroy(gqs_server* s);

/* ---- Device plumbing from the library's own HIP runtime -----------------
 * For callers behind a foreign-function layer.  Streams and events are
 * hipStream_t / hipEvent_t passed as void*. */
gqs_status gqs_malloc(gqs_ctx* ctx, size_t bytes, void** out);
gqs_status gqs_free(gqs_ctx* ctx, void* p);
gqs_status gqs_host_alloc(gqs_ctx* ctx, size_t bytes, void** out); /* pinned host memory */
gqs_status gqs_host_free(gqs_ctx* ctx, void* p);
gqs_status gqs_memcpy(gqs_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);       /* synchronous */
gqs_status gqs_memcpy_async(gqs_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream); /* stream-ordered */
gqs_status gqs_stream_create(gqs_ctx* ctx, void** out);
gqs_status gqs_stream_destroy(gqs_ctx* ctx, void* stream);
gqs_status gqs_stream_sync(gqs_ctx* ctx, void* stream);
gqs_status gqs_event_create(gqs_ctx* ctx, void** out);
gqs_status gqs_event_destroy(gqs_ctx* ctx, void* ev);
gqs_status gqs_event_record(gqs_ctx* ctx, void* ev, void* stream);
gqs_status gqs_event_elapsed(gqs_ctx* ctx, void* start, void* stop, float* ms); /* waits for stop */

#ifdef __cplusplus
}
#endif
#endif /* GQSORT_H */
