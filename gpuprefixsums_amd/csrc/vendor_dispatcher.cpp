// The vendor comparator class (include/gps_vendor.hpp) and its C entry
// point, built into libgps_vendor.so beside vendor_rocprim.hip.
// Reference: GPUPrefixSumsCUDA/CubDispatcher.cuh:21-76.
#include "../../include/gps_vendor.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

namespace gps {

namespace {

void check(int rc, const char* what) {
    if (rc != GPS_OK) throw Error(rc, std::string(what) + ": " + gps_last_error());
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(GPS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

CubDispatcher::CubDispatcher(uint32_t size) : k_maxSize(size) {
    check_hip(hipMalloc(&m_scan, (size_t)std::max(k_maxSize, 1u) * sizeof(uint32_t)), "hipMalloc");
}

CubDispatcher::~CubDispatcher() { (void)hipFree(m_scan); }

double CubDispatcher::BatchTimingCubChainedScan(uint32_t size, uint32_t batchCount) {
    if (size > k_maxSize || size == 0) {
        printf("Error, requested test size exceeds max initialized size. \n");
        return 0.0;
    }
    printf("Beginning rocPRIM inclusive scan batch timing test at:\n");
    printf("Size: %u\n", size);
    printf("Test size: %u\n", batchCount);
    uint64_t temp_bytes = 0;
    check(gps_vendor_scan_temp_bytes(size, 1, &temp_bytes), "rocPRIM temp query");
    // No throw between the allocations and their release: errors are kept in
    // `e` / `rc` and raised after the clean-up.
    void* temp = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    hipError_t e = hipMalloc(&temp, std::max<uint64_t>(temp_bytes, 4));
    if (e == hipSuccess) e = hipEventCreate(&start);
    if (e == hipSuccess) e = hipEventCreate(&stop);
    double total_ms = 0.0;
    int rc = GPS_OK;
    for (uint32_t i = 0; i <= batchCount && rc == GPS_OK && e == hipSuccess; ++i) {
        rc = gps_fill_u32(m_scan, size, 1u, nullptr);
        if (rc == GPS_OK) e = hipDeviceSynchronize();
        if (rc == GPS_OK && e == hipSuccess) e = hipEventRecord(start, nullptr);
        if (rc == GPS_OK && e == hipSuccess) rc = gps_vendor_scan_u32(m_scan, m_scan, size, 1, temp, temp_bytes, nullptr);
        if (rc == GPS_OK && e == hipSuccess) e = hipEventRecord(stop, nullptr);
        if (rc == GPS_OK && e == hipSuccess) e = hipEventSynchronize(stop);
        float ms = 0.f;
        if (rc == GPS_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, start, stop);
        if (i) total_ms += ms;
        if ((i & 15) == 0) printf(". ");
    }
    uint64_t errors = 0;
    if (rc == GPS_OK && e == hipSuccess) rc = gps_validate_ones_u32(m_scan, size, 1, nullptr, &errors);
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
    (void)hipFree(temp);
    check_hip(e, "rocPRIM batch timing");
    check(rc, "rocPRIM batch timing");
    if (errors) throw Error(GPS_ERR_DEVICE, "rocPRIM scan of all-ones input has mismatches");
    const double secs = total_ms / 1000.0;
    const double keys = secs > 0 ? (double)size * batchCount / secs : 0.0;
    printf("\nTotal time elapsed: %f\n", secs);
    printf("Estimated speed at %u 32-bit elements: %E keys/sec (%.1f GB/s at 8 B/key)\n\n", size, keys,
           keys * 8.0 / 1e9);
    return keys;
}

}  // namespace gps

extern "C" int gps_vendor_batch_timing(uint32_t max_size, uint32_t size, uint32_t batch_count,
                                       double* keys_per_sec) {
    if (!keys_per_sec) return GPS_ERR_INVALID;
    try {
        gps::CubDispatcher d(max_size);
        *keys_per_sec = d.BatchTimingCubChainedScan(size, batch_count);
    } catch (const gps::Error& e) {
        return e.code;
    }
    fflush(stdout);
    return GPS_OK;
}
