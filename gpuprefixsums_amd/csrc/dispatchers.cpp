// The reference's dispatcher classes (include/gps_dispatchers.hpp) over the C ABI.
#include "../../include/gps_dispatchers.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <vector>

namespace gps {

namespace {

void check(int rc, const char* what) {
    if (rc != GPS_OK) throw Error(rc, std::string(what) + ": " + gps_last_error());
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(GPS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

DispatcherBase::DispatcherBase(int algo, uint32_t maxSize, gps_tuning tuning)
    : k_maxSize((maxSize + 3u) & ~3u) {  // align16 (...Dispatcher.cuh:168): whole uint4s
    check(gps_create(algo, k_maxSize ? k_maxSize : 4u, &tuning, &m_handle), "gps_create");
    gps_tuning t;
    check(gps_get_tuning(m_handle, &t, &k_partitionSize), "gps_get_tuning");
    hipError_t e = hipMalloc(&m_scan, (size_t)std::max(k_maxSize, 4u) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&m_scratch, (size_t)std::max(k_maxSize, 4u) * sizeof(uint32_t));
    if (e != hipSuccess) {
        (void)hipFree(m_scan);
        gps_destroy(m_handle);
        check_hip(e, "hipMalloc");
    }
}

DispatcherBase::~DispatcherBase() {
    (void)hipFree(m_scan);
    (void)hipFree(m_scratch);
    gps_destroy(m_handle);
}

void DispatcherBase::Inclusive(const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream) {
    check(gps_inclusive_scan_u32(m_handle, d_in, d_out, n, stream), "inclusive scan");
}

void DispatcherBase::Exclusive(const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream) {
    check(gps_exclusive_scan_u32(m_handle, d_in, d_out, n, stream), "exclusive scan");
}

void DispatcherBase::TestAll(bool inclusive, const char* name) {
    printf("Beginning GPUPrefixSums %s validation test:\n", name);
    m_passed = m_total = 0;
    auto run_ones = [&](uint32_t n) {
        check(gps_fill_u32(m_scan, n, 1u, nullptr), "InitOne");
        if (inclusive)
            Inclusive(m_scan, m_scan, n);
        else
            Exclusive(m_scan, m_scan, n);
        uint64_t errors = 0;
        check(gps_validate_ones_u32(m_scan, n, inclusive ? 1 : 0, nullptr, &errors), "validate");
        ++m_total;
        if (errors == 0)
            ++m_passed;
        else
            printf("\n Test failed at size %u (%llu errors)\n", n, (unsigned long long)errors);
    };
    // Every size from one tile to two tiles (...Dispatcher.cuh:48,88).
    const uint32_t T = k_partitionSize;
    for (uint32_t n = T; n <= 2 * T && n <= k_maxSize; ++n) {
        run_ones(n);
        if (!(n & 255)) printf(".");
    }
    printf("\n");
    // Large sizes 2^26..2^28 (...Dispatcher.cuh:61-68, which passes i instead
    // of 1 << i to InitOne/Validate; corrected here), as far as maxSize allows.
    for (uint32_t i = 26; i <= 28; ++i)
        if ((1u << i) <= k_maxSize) run_ones(1u << i);
    // Full-range random input, checked with the parallel form of the serial
    // validator (GPUPrefixSumsD3D12/Shaders/Utility.hlsl:99-125), wrapping.
    const uint32_t nr = std::min<uint32_t>(k_maxSize, 1u << 20);
    check(gps_fill_random_u32(m_scratch, nr, 0x5EEDull + nr, 0, 32, nullptr), "InitRandom");
    if (inclusive)
        Inclusive(m_scratch, m_scan, nr);
    else
        Exclusive(m_scratch, m_scan, nr);
    uint64_t errors = 0;
    check(gps_validate_scan_u32(m_scratch, m_scan, nr, inclusive ? 1 : 0, nullptr, &errors), "validate");
    ++m_total;
    if (errors == 0)
        ++m_passed;
    else
        printf("\n Random test failed at size %u (%llu errors)\n", nr, (unsigned long long)errors);
    uint32_t dev_err = 0;
    check(gps_device_error(m_handle, &dev_err), "device error query");
    if (dev_err) {
        printf(" Device watchdog raised error bits 0x%x\n", dev_err);
        m_passed = 0;
    }
    if (m_passed == m_total)
        printf("%u/%u All tests passed.\n\n", m_passed, m_total);
    else
        printf("%u/%u Test failed.\n\n", m_passed, m_total);
}

double DispatcherBase::BatchTiming(bool inclusive, uint32_t size, uint32_t batchCount, const char* name) {
    if (size > k_maxSize) {
        printf("Error, requested test size exceeds max initialized size. \n");
        return 0.0;
    }
    printf("Beginning GPUPrefixSums %s batch timing test at:\n", name);
    printf("Size: %u\n", size);
    printf("Test size: %u\n", batchCount);
    hipEvent_t start, stop;
    check_hip(hipEventCreate(&start), "hipEventCreate");
    check_hip(hipEventCreate(&stop), "hipEventCreate");
    double total_ms = 0.0;
    // First iteration discarded as warm-up (...Dispatcher.cuh:134-149).
    for (uint32_t i = 0; i <= batchCount; ++i) {
        check(gps_fill_u32(m_scan, size, 1u, nullptr), "InitOne");
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check_hip(hipEventRecord(start, nullptr), "hipEventRecord");
        if (inclusive)
            Inclusive(m_scan, m_scan, size);
        else
            Exclusive(m_scan, m_scan, size);
        check_hip(hipEventRecord(stop, nullptr), "hipEventRecord");
        check_hip(hipEventSynchronize(stop), "hipEventSynchronize");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
        if (i) total_ms += ms;
        if ((i & 15) == 0) printf(". ");
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    const double secs = total_ms / 1000.0;
    const double keys = secs > 0 ? (double)size * batchCount / secs : 0.0;
    printf("\nTotal time elapsed: %f\n", secs);
    printf("Estimated speed at %u 32-bit elements: %E keys/sec (%.1f GB/s at 8 B/key)\n\n", size, keys,
           keys * 8.0 / 1e9);
    return keys;
}

EmulatedDeadlockingDispatcher::EmulatedDeadlockingDispatcher(uint32_t maxSize, uint32_t stallPeriod,
                                                             uint32_t maxSpinCount)
    : DispatcherBase(GPS_ALGO_CSDLDF, maxSize, gps_tuning{0, 0, maxSpinCount, stallPeriod}),
      m_stallPeriod(stallPeriod),
      m_spinLimit(maxSpinCount) {}

void EmulatedDeadlockingDispatcher::Rebuild(uint32_t stallPeriod, uint32_t spinLimit) {
    if (stallPeriod == m_stallPeriod && spinLimit == m_spinLimit) return;
    gps_tuning t{0, 0, spinLimit, stallPeriod};
    gps_handle* h = nullptr;
    check(gps_create(GPS_ALGO_CSDLDF, k_maxSize, &t, &h), "gps_create");
    gps_destroy(m_handle);
    m_handle = h;
    m_stallPeriod = stallPeriod;
    m_spinLimit = spinLimit;
}

double EmulatedDeadlockingDispatcher::BatchTimingInclusive(uint32_t size, uint32_t batchCount, uint32_t partMask,
                                                           uint32_t maxSpinCount) {
    Rebuild(partMask + 1u, maxSpinCount);
    return BatchTiming(true, size, batchCount, "EmulatedDeadlocking inclusive");
}

void EmulatedDeadlockingDispatcher::RecordTimingData(uint32_t spinCount) {
    std::ofstream writer("data.txt");
    writer << "Max Spin Count: " << spinCount << "\n";
    for (uint32_t i = 512; i >= 64; i >>= 1) {
        const uint32_t size = std::min<uint32_t>(1u << 28, k_maxSize);
        writer << BatchTimingInclusive(size, 250, i - 1, spinCount) << "\n";
    }
}

}  // namespace gps

// C entry point used by the test suite to run the dispatcher self-tests.
extern "C" int gps_dispatcher_test_all(int algo, uint32_t max_size, int inclusive, uint32_t* passed,
                                       uint32_t* total) {
    if (!passed || !total) return GPS_ERR_INVALID;
    try {
        if (algo == GPS_ALGO_CSDL) {
            gps::ChainedScanDecoupledLookbackDispatcher d(max_size);
            inclusive ? d.TestAllInclusive() : d.TestAllExclusive();
            *passed = d.TestsPassed();
            *total = d.TestsRun();
        } else if (algo == GPS_ALGO_RTS) {
            gps::ReduceThenScanDispatcher d(max_size);
            inclusive ? d.TestAllInclusive() : d.TestAllExclusive();
            *passed = d.TestsPassed();
            *total = d.TestsRun();
        } else if (algo == GPS_ALGO_CSDLDF) {
            gps::EmulatedDeadlockingDispatcher d(max_size);
            inclusive ? d.TestAllInclusive() : d.TestAllExclusive();
            *passed = d.TestsPassed();
            *total = d.TestsRun();
        } else {
            return GPS_ERR_INVALID;
        }
    } catch (const gps::Error& e) {
        return e.code;
    }
    fflush(stdout);
    return GPS_OK;
}
