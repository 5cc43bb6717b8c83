// generate.cpp -- synthetic inputs for the benchmark configurations.
//
// Every key is a pure function of (seed, index) through a splitmix64-style
// mixer, so any thread split gives the same array.  Generators:
//   0 uniform u32 over [0, 2^32)
//   1 "gaussian": floor of the mean of four independent uniform u32 draws
//   2 sorted: stratified uniform, strictly increasing
//   3 reverse sorted: generator 2 read backwards
//   4 all equal: 0x9E3779B9
//   5 bucket: 256 runs of n/256 keys, run r uniform in [p(r)*2^24, (p(r)+1)*2^24)
//     with p a seeded Fisher-Yates permutation of 0..255
//   6 permutation of 1..n, the reference's own data
//     (CL12/QuicksortMain.cpp:415-416; FSY/QuicksortMain.cpp:975-976), shuffled
//     with a seeded Fisher-Yates in place of std::random_shuffle
//   7 (f32/f64) N(0,1) by Box-Muller with u1 in (0,1]; 1% replaced by uniform
//     values in [-1e30, 1e30]; 0.1% replaced by a copy of another element;
//     -0 is written as +0; no NaN and no infinity can occur.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gqsort.h"

namespace {

inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed * 0xD1B54A32D192ED03ull + stream) ^ i);
}
inline uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
inline double unit53(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)

template <class F>
void parallel_for(size_t n, F f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (n < (1u << 16)) nt = 1;
    std::vector<std::thread> ws;
    for (unsigned t = 0; t < nt; ++t) {
        size_t lo = n * t / nt, hi = n * (t + 1) / nt;
        ws.emplace_back([=] {
            for (size_t i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto& w : ws) w.join();
}

uint32_t sorted_key(uint64_t seed, uint64_t i, uint64_t n) {
    const uint64_t lo = (i << 32) / n;
    const uint64_t hi = ((i + 1) << 32) / n;
    const uint64_t width = hi > lo ? hi - lo : 1;
    return (uint32_t)(lo + rnd(seed, 2, i) % width);
}

double normal_base(uint64_t seed, uint64_t i) {
    const double u1 = 1.0 - unit53(rnd(seed, 10, i));  // (0, 1]
    const double u2 = unit53(rnd(seed, 11, i));
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    if ((rnd(seed, 12, i) % 10000) < 100) z = (2.0 * unit53(rnd(seed, 13, i)) - 1.0) * 1e30;
    return z;
}

template <class T>
T float_key(uint64_t seed, uint64_t i, uint64_t n) {
    double v = normal_base(seed, i);
    if ((rnd(seed, 14, i) % 10000) < 10) v = normal_base(seed, rnd(seed, 15, i) % n);
    T t = (T)v;
    if (t == T(0)) t = T(0);  // canonical +0
    return t;
}

}  // namespace

extern "C" gqs_status gqs_generate(int dtype, int gen, uint64_t seed, void* out, size_t n) {
    if (!out && n) return GQS_EINVAL;
    if (n == 0) return GQS_OK;
    if (dtype == GQS_U32) {
        uint32_t* a = static_cast<uint32_t*>(out);
        if (n > 0xFFFFFFFFull && (gen == 2 || gen == 3 || gen == 6)) return GQS_EINVAL;
        switch (gen) {
            case 0: parallel_for(n, [=](size_t i) { a[i] = hi32(rnd(seed, 0, i)); }); break;
            case 1:
                parallel_for(n, [=](size_t i) {
                    uint64_t s = 0;
                    for (int k = 0; k < 4; ++k) s += hi32(rnd(seed, 1, 4 * (uint64_t)i + k));
                    a[i] = (uint32_t)(s / 4);
                });
                break;
            case 2: parallel_for(n, [=](size_t i) { a[i] = sorted_key(seed, i, n); }); break;
            case 3: parallel_for(n, [=](size_t i) { a[i] = sorted_key(seed, n - 1 - i, n); }); break;
            case 4: parallel_for(n, [=](size_t i) { a[i] = 0x9E3779B9u; }); break;
            case 5: {
                uint32_t perm[256];
                for (int b = 0; b < 256; ++b) perm[b] = (uint32_t)b;
                for (int b = 255; b > 0; --b) {
                    int j = (int)(rnd(seed, 5, (uint64_t)b) % (uint64_t)(b + 1));
                    uint32_t t = perm[b];
                    perm[b] = perm[j];
                    perm[j] = t;
                }
                std::vector<uint32_t> pv(perm, perm + 256);
                parallel_for(n, [=](size_t i) {
                    const uint64_t run = (uint64_t)i * 256 / n;
                    a[i] = (pv[run] << 24) | (hi32(rnd(seed, 6, i)) & 0xFFFFFFu);
                });
                break;
            }
            case 6: {
                for (size_t i = 0; i < n; ++i) a[i] = (uint32_t)(i + 1);
                for (size_t i = n - 1; i > 0; --i) {
                    size_t j = rnd(seed, 7, i) % (i + 1);
                    uint32_t t = a[i];
                    a[i] = a[j];
                    a[j] = t;
                }
                break;
            }
            default: return GQS_EINVAL;
        }
        return GQS_OK;
    }
    if (dtype == GQS_F32 || dtype == GQS_F64) {
        if (gen == 7) {
            if (dtype == GQS_F32) {
                float* a = static_cast<float*>(out);
                parallel_for(n, [=](size_t i) { a[i] = float_key<float>(seed, i, n); });
            } else {
                double* a = static_cast<double*>(out);
                parallel_for(n, [=](size_t i) { a[i] = float_key<double>(seed, i, n); });
            }
            return GQS_OK;
        }
        if (gen == 6) {
            // FSY's data: integer-valued keys 1..n shuffled (FSY/QuicksortMain.cpp:975-976).
            std::vector<uint64_t> idx(n);
            for (size_t i = 0; i < n; ++i) idx[i] = i + 1;
            for (size_t i = n - 1; i > 0; --i) {
                size_t j = rnd(seed, 7, i) % (i + 1);
                std::swap(idx[i], idx[j]);
            }
            if (dtype == GQS_F32) {
                float* a = static_cast<float*>(out);
                for (size_t i = 0; i < n; ++i) a[i] = (float)idx[i];
            } else {
                double* a = static_cast<double*>(out);
                for (size_t i = 0; i < n; ++i) a[i] = (double)idx[i];
            }
            return GQS_OK;
        }
        return GQS_EINVAL;
    }
    return GQS_EINVAL;
}
