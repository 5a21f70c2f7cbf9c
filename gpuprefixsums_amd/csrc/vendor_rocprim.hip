// Vendor comparator: rocPRIM's device-wide scan (rocprim::inclusive_scan /
// exclusive_scan, rocPRIM 4.2 in ROCm 7.2) behind the same kind of C entry
// points as the library's own scans (include/gps_vendor.h). Built into the
// comparator library libgps_vendor.so, never into libgps.so.
//
// Replaces the reference's CUB timing comparator
// (GPUPrefixSumsCUDA/CubDispatcher.cuh:32-76), which times
// cub::DeviceScan::InclusiveSum with a two-phase temporary-storage query. On
// ROCm the vendor's single-pass decoupled-lookback scan is rocPRIM's. Unlike
// the reference's comparator, this one takes caller buffers (in == out is
// allowed by rocPRIM) and its results are checked against the oracle by the
// test suite.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_scan.hpp>
#include <vector>

#include "../../include/gps_vendor.h"
#include "handle.hpp"

using gps::fail;
using gps::hip_fail;

namespace {

hipError_t vendor_scan(void* temp, size_t& bytes, const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive,
                       hipStream_t s) {
    if (inclusive) return rocprim::inclusive_scan(temp, bytes, in, out, (size_t)n, rocprim::plus<uint32_t>(), s);
    return rocprim::exclusive_scan(temp, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s);
}

int check_vendor_args(const uint32_t* in, uint32_t* out, uint64_t n, void* temp, uint64_t temp_bytes,
                      int inclusive) {
    if (n == 0) return GPS_OK;
    if (!in || !out) return fail(GPS_ERR_INVALID, "null buffer");
    size_t need = 0;
    hipError_t e = vendor_scan(nullptr, need, in, out, n, inclusive != 0, nullptr);
    if (e != hipSuccess) return hip_fail(e, "rocprim scan size query");
    if (need > 0 && (!temp || temp_bytes < need))
        return fail(GPS_ERR_INVALID, "rocPRIM scan needs %zu bytes of temporary storage, got %llu", need,
                    (unsigned long long)temp_bytes);
    return GPS_OK;
}

}  // namespace

extern "C" {

int gps_vendor_scan_temp_bytes(uint64_t n, int inclusive, uint64_t* bytes) {
    if (!bytes) return fail(GPS_ERR_INVALID, "null bytes pointer");
    size_t need = 0;
    if (n) {
        // The query reads no memory; any non-null pointers do.
        const uint32_t* dummy = reinterpret_cast<const uint32_t*>(uintptr_t(16));
        hipError_t e = vendor_scan(nullptr, need, dummy, const_cast<uint32_t*>(dummy), n, inclusive != 0, nullptr);
        if (e != hipSuccess) return hip_fail(e, "rocprim scan size query");
    }
    *bytes = need;
    return GPS_OK;
}

int gps_vendor_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* d_temp,
                        uint64_t temp_bytes, void* stream) {
    int rc = check_vendor_args(d_in, d_out, n, d_temp, temp_bytes, inclusive);
    if (rc || n == 0) return rc;
    size_t bytes = temp_bytes;
    HIP_TRY(vendor_scan(d_temp, bytes, d_in, d_out, n, inclusive != 0, static_cast<hipStream_t>(stream)));
    return GPS_OK;
}

int gps_vendor_time_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* d_temp,
                             uint64_t temp_bytes, void* stream, uint32_t iters, float* ms) {
    if (!ms || iters == 0 || n == 0) return fail(GPS_ERR_INVALID, "bad timing arguments");
    int rc = check_vendor_args(d_in, d_out, n, d_temp, temp_bytes, inclusive);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<hipEvent_t> ev(iters + 1);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    int status = GPS_OK;
    hipError_t e = hipEventRecord(ev[0], s);
    if (e != hipSuccess) status = hip_fail(e, "hipEventRecord");
    for (uint32_t i = 0; i < iters && status == GPS_OK; ++i) {
        size_t bytes = temp_bytes;
        e = vendor_scan(d_temp, bytes, d_in, d_out, n, inclusive != 0, s);
        if (e == hipSuccess) e = hipEventRecord(ev[i + 1], s);
        if (e != hipSuccess) status = hip_fail(e, "rocprim scan");
    }
    if (status == GPS_OK) {
        e = hipEventSynchronize(ev[iters]);
        if (e != hipSuccess) status = hip_fail(e, "hipEventSynchronize");
        for (uint32_t i = 0; i < iters && status == GPS_OK; ++i) {
            e = hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
            if (e != hipSuccess) status = hip_fail(e, "hipEventElapsedTime");
        }
    }
    for (auto& ev_i : ev) (void)hipEventDestroy(ev_i);
    return status;
}

}  // extern "C"
