// gqsort.hpp -- C++ drop-in for the reference's GPUQSort entry point.
//
// Replaces
//     template <class T> void GPUQSort(OCLResources* pOCL, size_t size, T* d, T* dn)
//     (GPU-Quicksort_OpenCL_1.2/QuicksortMain.cpp:297-298,
//      GPU-Quicksort_SYCL/QuicksortMain.cpp:357-358,
//      GPU-Quicksort_Full_SYCL/QuicksortMain.cpp:835-836)
// with the same argument meaning: d (host memory) holds `size` keys and is
// sorted ascending in place, dn is caller-owned scratch of `size` keys.  The
// OCLResources handle becomes a gqs_ctx*; the return value is a status instead
// of an assert (CL12/OpenCLUtils.h:41-45).  T is uint32_t (the reference's
// `uint`), float or double (FSY/QuicksortMain.cpp:1192-1194).
#ifndef GQSORT_HPP
#define GQSORT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>

#include "gqsort.h"

template <class T> struct gqs_dtype_of;
template <> struct gqs_dtype_of<uint32_t> { static constexpr int value = GQS_U32; };
template <> struct gqs_dtype_of<float> { static constexpr int value = GQS_F32; };
template <> struct gqs_dtype_of<double> { static constexpr int value = GQS_F64; };

// Host pointers, synchronous (the reference's semantics).
template <class T>
inline gqs_status GPUQSort(gqs_ctx* ctx, size_t size, T* d, T* dn) {
    return gqs_sort(ctx, gqs_dtype_of<T>::value, d, dn, size);
}

// Device pointers, stream-ordered.
template <class T>
inline gqs_status GPUQSortDevice(gqs_ctx* ctx, size_t size, T* d, T* dn, void* stream = nullptr) {
    return gqs_sort_device(ctx, gqs_dtype_of<T>::value, d, dn, size, stream);
}

// RAII context for C++ callers (the role of OCLResources + InitializeOpenCL,
// CL12/QuicksortMain.cpp:83-90, 465-475).
class GqsContext {
public:
    explicit GqsContext(int device = 0) {
        if (gqs_create(&ctx_, device) != GQS_OK) throw std::runtime_error("gqs_create failed");
    }
    ~GqsContext() { gqs_destroy(ctx_); }
    GqsContext(const GqsContext&) = delete;
    GqsContext& operator=(const GqsContext&) = delete;
    gqs_ctx* get() const { return ctx_; }
    operator gqs_ctx*() const { return ctx_; }
    std::string last_error() const { return gqs_last_error(ctx_); }

private:
    gqs_ctx* ctx_ = nullptr;
};

#endif  // GQSORT_HPP
