// keys.h -- device-side key traits shared by the sort and merge kernels.
//
// Keys are compared as unsigned integers: float/double bits go through an
// order-preserving bijection (sign bit set for non-negative values, all bits
// flipped for negative ones) on load and back on store.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "records.h"

namespace gqs {

template <int DT> struct KeyOf { using K = uint32_t; };
template <> struct KeyOf<DT_F64> { using K = uint64_t; };

template <int DT>
__device__ __forceinline__ typename KeyOf<DT>::K to_key(typename KeyOf<DT>::K u) {
    if constexpr (DT == DT_F32) return (u >> 31) ? ~u : (u | 0x80000000u);
    else if constexpr (DT == DT_F64) return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
    else return u;
}
template <int DT>
__device__ __forceinline__ typename KeyOf<DT>::K from_key(typename KeyOf<DT>::K k) {
    if constexpr (DT == DT_F32) return (k >> 31) ? (k & 0x7FFFFFFFu) : ~k;
    else if constexpr (DT == DT_F64) return (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    else return k;
}
template <class K>
__device__ __forceinline__ K key_max() { return ~K(0); }

template <int DT>
__device__ __forceinline__ typename KeyOf<DT>::K load_key(const typename KeyOf<DT>::K* p, uint64_t i, bool raw) {
    typename KeyOf<DT>::K v = p[i];
    return raw ? to_key<DT>(v) : v;
}

}  // namespace gqs
