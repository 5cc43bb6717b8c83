// quicksort_main.cpp -- command-line driver with the reference's argv and
// output lines, on top of libgqsort's C ABI.
//
//   quicksort <iterations> <cpu|gpu> <intel|amd|nvidia> <width> <height> <show_CL|no_show_CL> [uint|float|double|all]
//
// The first six arguments are the reference's (CL12/QuicksortMain.cpp:112-136,
// FSY/QuicksortMain.cpp:140-165): the array holds width*height keys, a shuffled
// permutation of 1..width*height (CL12/QuicksortMain.cpp:415-416;
// FSY/QuicksortMain.cpp:975-976).  The optional seventh picks the key type;
// "all" runs uint, float and double in turn as FSY's main does
// (FSY/QuicksortMain.cpp:1192-1194).  The printed lines follow the reference's
// (CL12/QuicksortMain.cpp:418-540): CPU std::sort and parallel sort times,
// "verifying: true|false", one "Time to sort" per GPUQSort call, the number of
// failures, average, standard deviation, min/max and the speedups over the CPU
// sorts.  Each GPUQSort call takes host pointers and so includes the PCIe
// transfers, as the reference's host-pointer call does.
//
// Differences: the device is the first HIP GPU whatever the vendor string
// names ("cpu" is refused: this library has no CPU device); the shuffle is
// std::shuffle with a fixed-seed std::mt19937 (std::random_shuffle is gone
// from C++17), so the permutation differs from the reference's; the parallel
// CPU sort is a chunked std::sort plus pairwise merges on std::thread (TBB is
// not required); "Time to create the GPU context" replaces the OpenCL program
// build time.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include <unistd.h>

#include "gqsort.h"

namespace {

double seconds() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

const char kUsage[] =
    "Usage: Quicksort [num test iterations] [cpu|gpu] [intel|amd|nvidia] [SurfWidth(^2 only)] "
    "[SurfHeight(^2 only)] [show_CL | no_show_CL] [uint|float|double|all]";

// Sorts with `threads` std::threads: chunks sorted with std::sort, then merged
// pairwise (the role tbb::parallel_sort plays in the reference).
template <class T>
void parallel_sort(T* a, size_t n, unsigned threads) {
    if (threads < 2 || n < (size_t)threads * 4096) {
        std::sort(a, a + n);
        return;
    }
    std::vector<size_t> cut(threads + 1);
    for (unsigned i = 0; i <= threads; ++i) cut[i] = n * i / threads;
    std::vector<std::thread> pool;
    for (unsigned i = 0; i < threads; ++i) pool.emplace_back([&, i] { std::sort(a + cut[i], a + cut[i + 1]); });
    for (auto& t : pool) t.join();
    for (size_t width = 1; width < threads; width *= 2) {
        pool.clear();
        for (size_t i = 0; i + width < threads; i += 2 * width) {
            const size_t lo = cut[i], mid = cut[i + width], hi = cut[std::min<size_t>(i + 2 * width, threads)];
            pool.emplace_back([=] { std::inplace_merge(a + lo, a + mid, a + hi); });
        }
        for (auto& t : pool) t.join();
    }
}

// "verifying: true|false" against std::sort of the original keys.
template <class T>
bool verify(const std::vector<T>& original, const T* result, size_t n) {
    std::vector<T> expect(original);
    std::cout << "verifying: ";
    std::sort(expect.begin(), expect.end());
    const bool correct = std::equal(expect.begin(), expect.end(), result);
    std::cout << std::boolalpha << correct << std::endl;
    if (!correct) {
        size_t bad = 0;
        for (size_t i = 0; i < n; ++i) bad += expect[i] != result[i];
        std::cout << "num_discrepancies: " << bad << std::endl;
    }
    return correct;
}

template <class T>
int dtype_of() {
    if (std::is_same<T, float>::value) return GQS_F32;
    if (std::is_same<T, double>::value) return GQS_F64;
    return GQS_U32;
}

template <class T>
int big_test(gqs_ctx* ctx, size_t n, unsigned iterations, const std::string& device, const char* type_name) {
    printf("\n\n\n--------------------------------------------------------------------\n");
    printf("Allocating array size of %zu\n", n);
    std::vector<T> original(n), work(n), scratch(n);
    for (size_t i = 0; i < n; ++i) original[i] = (T)(i + 1);
    std::mt19937 rng;
    std::shuffle(original.begin(), original.end(), rng);

    std::cout << "Sorting the regular way..." << std::endl;
    work = original;
    double t0 = seconds();
    std::sort(work.begin(), work.end());
    const double std_sort_time = seconds() - t0;
    std::cout << "Time to sort: " << std_sort_time * 1000 << " ms" << std::endl;

    const unsigned threads = std::max(1u, std::thread::hardware_concurrency());
    std::cout << "Sorting with parallel quicksort on the cpu: " << std::endl;
    work = original;
    t0 = seconds();
    parallel_sort(work.data(), n, threads);
    const double par_sort_time = seconds() - t0;
    std::cout << "Time to sort: " << par_sort_time * 1000 << " ms" << std::endl;
    verify(original, work.data(), n);

    std::cout << "Sorting with GPUQSort on the " << device << " with type " << type_name << ": " << std::endl;
    std::vector<double> times(iterations);
    double average = 0.0;
    unsigned failures = 0;
    for (unsigned k = 0; k < iterations; ++k) {
        work = original;
        t0 = seconds();
        const gqs_status st = gqs_sort(ctx, dtype_of<T>(), work.data(), scratch.data(), n);
        const double t = seconds() - t0;
        if (st != GQS_OK) {
            std::cerr << "GPUQSort failed: " << gqs_last_error(ctx) << std::endl;
            return 1;
        }
        std::cout << "Time to sort: " << t * 1000 << " ms" << std::endl;
        times[k] = t;
        average += t;
        if (!verify(original, work.data(), n)) ++failures;
    }
    std::cout << " Number of failures: " << failures << " out of " << iterations << std::endl;
    average /= iterations;
    std::cout << "Average Time: " << average * 1000 << " ms" << std::endl;
    if (iterations > 1) {
        double dev = 0.0, lo = times[0], hi = times[0];
        for (double t : times) {
            dev += (average - t) * (average - t);
            lo = std::min(lo, t);
            hi = std::max(hi, t);
        }
        dev = std::sqrt(dev / (iterations - 1));
        std::cout << "Standard Deviation: " << dev * 1000 << std::endl;
        std::cout << "%error (3*stdDev)/Average: " << 3 * dev / average * 100 << "%" << std::endl;
        std::cout << "min time: " << lo * 1000 << " ms" << std::endl;
        std::cout << "max time: " << hi * 1000 << " ms" << std::endl;
    }
    std::cout << "Average speedup over CPU quicksort: " << par_sort_time / average << std::endl;
    std::cout << "Average speedup over CPU std::sort: " << std_sort_time / average << std::endl;
    printf("-------done--------------------------------------------------------\n");
    return failures ? 2 : 0;
}

// The reference's show_CL report (QueryPrintOpenCLDeviceInfo,
// CL12/OpenCLUtils.h:291-387): one line per property, for the HIP device.
void show_device(gqs_ctx* ctx) {
    gqs_device_props p{};
    if (gqs_device_report(ctx, &p) != GQS_OK) return;
    printf("Using platform: AMD ROCm/HIP and device: %s.\n", p.name);
    printf("HIP Device info:\n");
    printf("DEVICE_NAME                 :%s\n", p.name);
    printf("DEVICE_ARCH                 :%s\n", p.arch);
    printf("PCI_BUS_ID                  :%s\n", p.pci_bus_id);
    printf("DRIVER_VERSION              :%d\n", p.driver_version);
    printf("RUNTIME_VERSION             :%d\n", p.runtime_version);
    printf("MAX_COMPUTE_UNITS           :%8d\n", p.compute_units);
    printf("WAVEFRONT_SIZE              :%8d\n", p.wave_size);
    printf("MAX_WORK_ITEM_SIZES         :    (%5d, %5d, %5d)\n", p.max_block_dim[0], p.max_block_dim[1],
           p.max_block_dim[2]);
    printf("MAX_WORK_GROUP_SIZE         :%8d\n", p.max_threads_per_block);
    printf("MAX_CLOCK_FREQUENCY (MHz)   :%8d\n", p.clock_khz / 1000);
    printf("MEMORY_CLOCK (MHz)          :%8d\n", p.mem_clock_khz / 1000);
    printf("MEMORY_BUS_WIDTH (bits)     :%8d\n", p.mem_bus_bits);
    printf("LOCAL_MEM_SIZE (per WG)     :%12.1f\n", (double)p.lds_per_block);
    printf("LOCAL_MEM_SIZE (per CU)     :%12.1f\n", (double)p.lds_per_cu);
    printf("L2_CACHE_SIZE               :%12.1f\n", (double)p.l2_bytes);
    printf("GLOBAL_MEM_SIZE             :%12.1f\n", (double)p.global_mem_bytes);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 7 && argc != 8) {
        printf("%s\n", kUsage);
        return -1;
    }
    const int iterations = atoi(argv[1]);
    const std::string device = argv[2], vendor = argv[3];
    const long width = atol(argv[4]), height = atol(argv[5]);
    const bool show = argv[6][0] == 's';
    const std::string types = argc == 8 ? argv[7] : "uint";
    if (iterations < 1 || width < 1 || height < 1) {
        printf("%s\n", kUsage);
        return -1;
    }
    if (device != "gpu") {
        fprintf(stderr, "device '%s' is not available: GPUQSort runs on the GPU (use 'gpu')\n", device.c_str());
        return 1;
    }
    if (types != "uint" && types != "float" && types != "double" && types != "all") {
        printf("%s\n", kUsage);
        return -1;
    }
    const size_t n = (size_t)width * (size_t)height;
    const double t0 = seconds();
    gqs_ctx* ctx = nullptr;
    if (gqs_create(&ctx, 0) != GQS_OK) {
        fprintf(stderr, "no GPU: %s\n", ctx ? gqs_last_error(ctx) : "gqs_create failed");
        return 1;
    }
    // the tuned parameter table shipped next to this executable, when present
    {
        char exe[4096];
        const ssize_t len = readlink("/proc/self/exe", exe, sizeof exe - 1);
        if (len > 0) {
            exe[len] = 0;
            std::string table(exe);
            table = table.substr(0, table.rfind('/') + 1) + "tuned_mi355x.json";
            if (FILE* f = fopen(table.c_str(), "r")) {
                fclose(f);
                gqs_tune_load(ctx, table.c_str());
            }
        }
    }
    std::cout << "Time to create the GPU context: " << (seconds() - t0) * 1000 << " ms" << std::endl;
    if (vendor != "amd") printf("vendor '%s' requested; running on HIP device 0\n", vendor.c_str());
    if (show) show_device(ctx);
    int rc = 0;
    if (types == "uint" || types == "all") rc |= big_test<uint32_t>(ctx, n, (unsigned)iterations, device, "uint");
    if (types == "float" || types == "all") rc |= big_test<float>(ctx, n, (unsigned)iterations, device, "float");
    if (types == "double" || types == "all") rc |= big_test<double>(ctx, n, (unsigned)iterations, device, "double");
    gqs_destroy(ctx);
    return rc;
}
