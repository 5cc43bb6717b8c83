// ref_driver.cpp -- exports a few symbols of the reference's own shared header
// (GPU-Quicksort_OpenCL_1.2/Quicksort.h, compiled with HOST defined exactly as
// CL12/QuicksortMain.cpp:150-151 does) so that tests can pin the oracle's
// restatement against the real thing.  TEST INFRASTRUCTURE ONLY; built into
// oracle/_ref/ from the reference sources where they lie, never shipped.
#include <sys/types.h>
#include <cstddef>
#include <cstdint>
#define HOST 1
#include "Quicksort.h"

extern "C" {
// Host median-of-3 (CL12/Quicksort.h:34-57).
uint32_t ref_host_median_u32(uint32_t a, uint32_t b, uint32_t c) { return median<uint>(a, b, c); }
double ref_host_median_f64(double a, double b, double c) { return median<double>(a, b, c); }
// sizeof work_record / parent_record / block_record (CL12/Quicksort.h:105-145).
void ref_record_sizes(size_t* out) {
    out[0] = sizeof(work_record);
    out[1] = sizeof(parent_record);
    out[2] = sizeof(block_record);
}
// Default (Intel GPU) tunables and EMPTY_RECORD (CL12/Quicksort.h:91-100).
void ref_tunables(int* out) {
    out[0] = QUICKSORT_BLOCK_SIZE;
    out[1] = GQSORT_LOCAL_WORKGROUP_SIZE;
    out[2] = LQSORT_LOCAL_WORKGROUP_SIZE;
    out[3] = SORT_THRESHOLD;
    out[4] = EMPTY_RECORD;
}
}
