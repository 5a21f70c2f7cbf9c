/* gps_vendor.h - the vendor comparator: rocPRIM's device-wide scan behind C
 * entry points, in its own library (libgps_vendor.so, which links libgps.so
 * for its fill and validation kernels). It is a benchmark and test
 * comparator only: the product library libgps.so carries no vendor scan.
 *
 * Replaces the reference's CUB comparator (GPUPrefixSumsCUDA/CubDispatcher.cuh:
 * 32-76: cub::DeviceScan::InclusiveSum with a two-phase temporary-storage
 * query). On ROCm the vendor scan is rocprim::inclusive_scan / exclusive_scan.
 * Return codes are gps_status values; messages are read with gps_last_error()
 * from libgps.so. */
#ifndef GPS_VENDOR_H
#define GPS_VENDOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Caller buffers (d_in == d_out allowed), caller temporary storage of at
 * least gps_vendor_scan_temp_bytes(n, inclusive) bytes, no handle. */
int gps_vendor_scan_temp_bytes(uint64_t n, int inclusive, uint64_t* bytes);
int gps_vendor_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* d_temp,
                        uint64_t temp_bytes, void* stream);
/* `iters` back-to-back vendor scans, per-launch hipEvent times into ms[iters]
 * (the timing loop of CubDispatcher.cuh:52-68). */
int gps_vendor_time_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* d_temp,
                             uint64_t temp_bytes, void* stream, uint32_t iters, float* ms);
/* Runs CubDispatcher(max_size).BatchTimingCubChainedScan(size, batch_count)
 * (gps_vendor.hpp) and writes its keys/s (CubDispatcher.cuh:32-76). */
int gps_vendor_batch_timing(uint32_t max_size, uint32_t size, uint32_t batch_count, double* keys_per_sec);

#ifdef __cplusplus
}
#endif

#endif /* GPS_VENDOR_H */
