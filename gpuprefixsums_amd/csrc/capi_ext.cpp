// C ABI of the scan forms beside the plain uint32 scan: batched segmented
// scan (one launch, graph-capturable) and its per-segment-launch baseline.
#include <hip/hip_runtime.h>

#include <new>

#include "handle.hpp"

using gps::fail;
using gps::hip_fail;

namespace {

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// The persistent segmented kernel unless the handle asks for one workgroup
// per tile.
bool seg_persistent(const gps_handle* h) { return h->schedule != GPS_SCHEDULE_TILE; }

// Segmented kernels use the handle's tile shape when it is one of theirs,
// otherwise 1024 threads x 8 uint4 (32768-element tiles). A default handle
// picks the shape from n, from the exhaustive sweep's timings on MI355X
// (mean time over the best shape, batched u32 workloads): 1024 x 2 below
// 2^20 elements (1.00-1.14 against 1.4-1.8 for 1024 x 8), 1024 x 4 below
// 2^22, 1024 x 8 from there (1.00 against 1.08-1.19 for 1024 x 4).
gps::TileShape seg_shape(const gps_handle* h, uint64_t n) {
    if (h->size_adaptive)
        return n < (1ull << 20) ? gps::TileShape{1024, 2} : n < (1ull << 22) ? gps::TileShape{1024, 4}
                                                                            : gps::TileShape{1024, 8};
    return gps::seg_shape_supported(h->shape, seg_persistent(h)) ? h->shape : gps::TileShape{1024, 8};
}

// FP kernels likewise: the handle's shape when it is an FP shape, else
// 1024 x 4. A default handle picks from n (sweep timings and back-to-back
// launches on MI355X): one FP scan uses 256 x 4 up to 2^18 elements (2^16:
// 4.8 against 9.0 us for 1024 x 4 ... 1024 x 8), 1024 x 4 up to 2^22 and
// 1024 x 8 above; the segmented FP kernel uses 256 x 4 below 2^20 and
// 1024 x 4 above (1.00 against 1.03-1.08 for 1024 x 8).
gps::TileShape fp_shape(const gps_handle* h, uint64_t n, bool segmented) {
    if (h->size_adaptive) {
        if (segmented) return n < (1ull << 20) ? gps::TileShape{256, 4} : gps::TileShape{1024, 4};
        return n <= (1ull << 18) ? gps::TileShape{256, 4} : n <= (1ull << 22) ? gps::TileShape{1024, 4}
                                                                               : gps::TileShape{1024, 8};
    }
    return gps::fp_shape_supported(h->shape) ? h->shape : gps::TileShape{1024, 4};
}

int check_seg(gps_handle* h, const void* in, const void* out, const void* offsets, uint64_t n) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    if (h->algo != GPS_ALGO_CSDL) return fail(GPS_ERR_UNSUPPORTED, "segmented scans need a CSDL handle");
    if (n && (!in || !out)) return fail(GPS_ERR_INVALID, "null data pointer");
    if (!offsets) return fail(GPS_ERR_INVALID, "null offsets");
    if (n > h->max_n)
        return fail(GPS_ERR_TOO_LARGE, "n=%llu exceeds handle max_n=%llu", (unsigned long long)n,
                    (unsigned long long)h->max_n);
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 3u)
        return fail(GPS_ERR_INVALID, "uint32 buffers must be 4-byte aligned");
    if (reinterpret_cast<uintptr_t>(offsets) & 7u) return fail(GPS_ERR_INVALID, "offsets must be 8-byte aligned");
    return GPS_OK;
}

int enqueue_segmented(gps_handle* h, const uint32_t* in, uint32_t* out, const uint64_t* offsets,
                      uint64_t num_segments, uint64_t n, bool inclusive, hipStream_t s) {
    if (n == 0) return GPS_OK;
    HIP_TRY(gps::before_epoch_launch(h, s));
    HIP_TRY(gps::launch_segscan_u32(seg_shape(h, n), inclusive, seg_persistent(h), in, out, offsets, num_segments, n,
                                    aligned16(in) && aligned16(out), h->status, h->ticket, h->err, s));
    return GPS_OK;
}

}  // namespace

namespace {

// One FP scan with pair or plain carries runs the persistent kernel on a
// handle with the persistent schedule (the CSDL default); a default handle
// picks it above 2^24 elements (2^25 FP32 pairs: 4398 against 4319 GB/s for
// one workgroup per tile; 2^23: 3663 against 3778).
bool fp_persistent(const gps_handle* h, uint64_t n) {
    if (h->size_adaptive) return n > (1ull << 24);
    return h->schedule == GPS_SCHEDULE_PERSISTENT;
}

int scan_fp(gps_handle* h, const void* in, float* out, uint64_t n, bool bf16, int compensated, void* stream) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    if (h->algo != GPS_ALGO_CSDL) return fail(GPS_ERR_UNSUPPORTED, "floating-point scans need a CSDL handle");
    if (n && (!in || !out)) return fail(GPS_ERR_INVALID, "null data pointer");
    if (n > h->max_n)
        return fail(GPS_ERR_TOO_LARGE, "n=%llu exceeds handle max_n=%llu", (unsigned long long)n,
                    (unsigned long long)h->max_n);
    if (compensated < gps::kFpPlain || compensated > gps::kFpPair)
        return fail(GPS_ERR_INVALID, "compensated must be 0 (plain) or 1 (two-float pairs)");
    const uintptr_t ia = reinterpret_cast<uintptr_t>(in), oa = reinterpret_cast<uintptr_t>(out);
    if ((ia & (bf16 ? 1u : 3u)) || (oa & 3u)) return fail(GPS_ERR_INVALID, "misaligned buffer");
    if (n == 0) return GPS_OK;
    const gps::TileShape s = fp_shape(h, n, false);
    const uint32_t vec_ok = ((ia & (bf16 ? 7u : 15u)) == 0) && ((oa & 15u) == 0);
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(gps::before_epoch_launch(h, st));
    HIP_TRY(gps::launch_scan_fp(s, bf16, compensated, fp_persistent(h, n), in, out, n, vec_ok, h->status, h->ticket,
                                h->err, st));
    return GPS_OK;
}

// The segmented FP scan runs its persistent kernel on a handle with the
// persistent schedule. A default handle runs it, at 1024 x 8, above 2^22
// elements (4096 log-normal segments of 50.9 M elements: FP32 pairs 0.1021
// against 0.1145 ms for one workgroup per tile at 1024 x 4, plain 0.0804
// against 0.1030 ms, BF16 pairs 0.0999 against 0.1032 ms).
bool segfp_persistent(const gps_handle* h, uint64_t n) {
    if (h->size_adaptive) return n > (1ull << 22);
    return h->schedule == GPS_SCHEDULE_PERSISTENT;
}

int segscan_fp(gps_handle* h, const void* in, float* out, const uint64_t* offsets, uint64_t num_segments,
               uint64_t n, bool bf16, int compensated, void* stream) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    if (h->algo != GPS_ALGO_CSDL) return fail(GPS_ERR_UNSUPPORTED, "floating-point scans need a CSDL handle");
    if (n && (!in || !out)) return fail(GPS_ERR_INVALID, "null data pointer");
    if (!offsets) return fail(GPS_ERR_INVALID, "null offsets");
    if (n > h->max_n)
        return fail(GPS_ERR_TOO_LARGE, "n=%llu exceeds handle max_n=%llu", (unsigned long long)n,
                    (unsigned long long)h->max_n);
    const uintptr_t ia = reinterpret_cast<uintptr_t>(in), oa = reinterpret_cast<uintptr_t>(out);
    if ((ia & (bf16 ? 1u : 3u)) || (oa & 3u)) return fail(GPS_ERR_INVALID, "misaligned buffer");
    if (reinterpret_cast<uintptr_t>(offsets) & 7u) return fail(GPS_ERR_INVALID, "offsets must be 8-byte aligned");
    if (compensated < gps::kFpPlain || compensated > gps::kFpPair)
        return fail(GPS_ERR_INVALID, "compensated must be 0 (plain) or 1 (two-float pairs)");
    if (n == 0) return GPS_OK;
    const bool persistent = segfp_persistent(h, n);
    const gps::TileShape s = persistent && h->size_adaptive ? gps::TileShape{1024, 8} : fp_shape(h, n, true);
    const uint32_t vec_ok = ((ia & (bf16 ? 7u : 15u)) == 0) && ((oa & 15u) == 0);
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(gps::before_epoch_launch(h, st));
    HIP_TRY(gps::launch_segscan_fp(s, bf16, compensated != 0, persistent, in, out, offsets, num_segments, n, vec_ok,
                                   h->status, h->ticket, h->err, st));
    return GPS_OK;
}

int scan_u32x4(gps_handle* h, const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive, void* stream) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    if (h->algo != GPS_ALGO_CSDL) return fail(GPS_ERR_UNSUPPORTED, "struct scans need a CSDL handle");
    if (n && (!in || !out)) return fail(GPS_ERR_INVALID, "null data pointer");
    if (n > h->max_n)
        return fail(GPS_ERR_TOO_LARGE, "n=%llu exceeds handle max_n=%llu", (unsigned long long)n,
                    (unsigned long long)h->max_n);
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u)
        return fail(GPS_ERR_INVALID, "uint4 struct buffers must be 16-byte aligned");
    if (n == 0) return GPS_OK;
    gps::TileShape s = gps::v4_shape_supported(h->shape) ? h->shape : gps::TileShape{1024, 8};
    bool persistent = h->schedule != GPS_SCHEDULE_TILE;
    if (h->size_adaptive) {
        // As for the plain scan, measured on MI355X: one workgroup per tile
        // below 2^22 structs, 1024 x 4 up to 2^19 (2^16 structs: 5.4 against
        // 6.7 us for the persistent 1024 x 8 kernel).
        if (n <= (1ull << 19)) {
            s = {1024, 4};
            persistent = false;
        } else if (n < (1ull << 22)) {
            s = {1024, 8};
            persistent = false;
        }
    }
    const uint64_t tile = (uint64_t)s.block * s.vec;
    if ((n + tile - 1) / tile > h->max_tiles) return fail(GPS_ERR_TOO_LARGE, "not enough status slots");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(gps::before_epoch_launch(h, st));
    HIP_TRY(gps::launch_csdl_u32x4(s, inclusive, persistent, in, out, n, h->status, h->ticket, h->err, st));
    return GPS_OK;
}

}  // namespace

struct gps_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int gps_segmented_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, const uint64_t* d_offsets,
                           uint64_t num_segments, uint64_t n, int inclusive, void* stream) {
    int rc = check_seg(h, d_in, d_out, d_offsets, n);
    if (rc) return rc;
    return enqueue_segmented(h, d_in, d_out, d_offsets, num_segments, n, inclusive != 0,
                             static_cast<hipStream_t>(stream));
}

int gps_segmented_scan_u32_per_segment(gps_handle* h, const uint32_t* d_in, uint32_t* d_out,
                                       const uint64_t* h_offsets, uint64_t num_segments, int inclusive,
                                       void* stream) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    if (!h_offsets) return fail(GPS_ERR_INVALID, "null offsets");
    for (uint64_t s = 0; s < num_segments; ++s) {
        const uint64_t a = h_offsets[s], b = h_offsets[s + 1];
        if (b < a) return fail(GPS_ERR_INVALID, "offsets must be non-decreasing (segment %llu)", (unsigned long long)s);
        if (b == a) continue;
        const int rc = inclusive ? gps_inclusive_scan_u32(h, d_in + a, d_out + a, b - a, stream)
                                 : gps_exclusive_scan_u32(h, d_in + a, d_out + a, b - a, stream);
        if (rc) return rc;
    }
    return GPS_OK;
}

int gps_segmented_graph_create(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, const uint64_t* d_offsets,
                               uint64_t num_segments, uint64_t n, int inclusive, gps_graph** out) {
    if (!out) return fail(GPS_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int rc = check_seg(h, d_in, d_out, d_offsets, n);
    if (rc) return rc;
    HIP_TRY(gps::prepare_segscan_u32(seg_shape(h, n), inclusive != 0, seg_persistent(h)));
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    gps_graph* g = new (std::nothrow) gps_graph();
    if (!g) {
        (void)hipStreamDestroy(s);
        return fail(GPS_ERR_INVALID, "out of host memory");
    }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        // No re-arm check inside a capture: a graph node must not depend on
        // the host-side launch count (gps_reset re-arms explicitly).
        rc = n ? (gps::launch_segscan_u32(seg_shape(h, n), inclusive != 0, seg_persistent(h), d_in, d_out, d_offsets,
                                          num_segments, n, aligned16(d_in) && aligned16(d_out), h->status,
                                          h->ticket, h->err, s) == hipSuccess
                      ? GPS_OK
                      : GPS_ERR_HIP)
               : GPS_OK;
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(s, &graph);
        g->graph = graph;
        if (rc != GPS_OK && e == hipSuccess) e = hipErrorLaunchFailure;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return hip_fail(e, "segmented graph capture");
    }
    *out = g;
    return GPS_OK;
}

int gps_inclusive_scan_f32(gps_handle* h, const float* d_in, float* d_out, uint64_t n, int compensated,
                           void* stream) {
    return scan_fp(h, d_in, d_out, n, false, compensated, stream);
}

int gps_inclusive_scan_bf16_f32(gps_handle* h, const uint16_t* d_in, float* d_out, uint64_t n, int compensated,
                                void* stream) {
    return scan_fp(h, d_in, d_out, n, true, compensated, stream);
}

int gps_segmented_scan_f32(gps_handle* h, const float* d_in, float* d_out, const uint64_t* d_offsets,
                           uint64_t num_segments, uint64_t n, int compensated, void* stream) {
    return segscan_fp(h, d_in, d_out, d_offsets, num_segments, n, false, compensated, stream);
}

int gps_segmented_scan_bf16_f32(gps_handle* h, const uint16_t* d_in, float* d_out, const uint64_t* d_offsets,
                                uint64_t num_segments, uint64_t n, int compensated, void* stream) {
    return segscan_fp(h, d_in, d_out, d_offsets, num_segments, n, true, compensated, stream);
}

int gps_survey_segment_u32(int alg, uint32_t* segment) {
    if (!segment) return fail(GPS_ERR_INVALID, "null argument");
    const uint32_t s = gps::survey_segment(alg);
    if (!s) return fail(GPS_ERR_INVALID, "unknown survey network %d", alg);
    *segment = s;
    return GPS_OK;
}

int gps_survey_scan_u32(int alg, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* stream) {
    const uint32_t seg = gps::survey_segment(alg);
    if (!seg) return fail(GPS_ERR_INVALID, "unknown survey network %d", alg);
    if (n % seg) return fail(GPS_ERR_INVALID, "n=%llu is not a multiple of the segment (%u)", (unsigned long long)n, seg);
    if (n == 0) return GPS_OK;
    if (!d_in || !d_out) return fail(GPS_ERR_INVALID, "null data pointer");
    const uintptr_t mask = alg == gps::kSurveyTileRegisters ? 15u : 3u;
    if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & mask)
        return fail(GPS_ERR_INVALID, "buffers must be %u-byte aligned", (unsigned)mask + 1u);
    HIP_TRY(gps::launch_survey_u32(alg, inclusive != 0, d_in, d_out, n, static_cast<hipStream_t>(stream)));
    return GPS_OK;
}

int gps_estimate_occupancy(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int struct_scan,
                           void* stream, uint32_t* measured, uint32_t* resident) {
    if (!h || !measured || !resident) return fail(GPS_ERR_INVALID, "null argument");
    if (int rc_ = gps::check_device(h)) return rc_;
    if (h->algo != GPS_ALGO_CSDL) return fail(GPS_ERR_UNSUPPORTED, "the occupancy estimate needs a CSDL handle");
    if (!d_in || !d_out) return fail(GPS_ERR_INVALID, "null data pointer");
    if (n == 0) return fail(GPS_ERR_INVALID, "n must be positive");
    if (n > h->max_n)
        return fail(GPS_ERR_TOO_LARGE, "n=%llu exceeds handle max_n=%llu", (unsigned long long)n,
                    (unsigned long long)h->max_n);
    const uintptr_t mask = struct_scan ? 15u : 3u;
    if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & mask)
        return fail(GPS_ERR_INVALID, struct_scan ? "uint4 struct buffers must be 16-byte aligned"
                                                 : "uint32 buffers must be 4-byte aligned");
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned long long* count = h->counters + gps::kCounterOccupancy;
    HIP_TRY(hipMemsetAsync(count, 0, sizeof(unsigned long long), st));
    HIP_TRY(gps::before_epoch_launch(h, st));
    if (struct_scan) {
        const gps::TileShape s = gps::v4_shape_supported(h->shape) ? h->shape : gps::TileShape{1024, 8};
        *resident = gps::v4_resident_workgroups(s, true);
        HIP_TRY(gps::launch_csdl_u32x4(s, true, true, d_in, d_out, n, h->status, h->ticket, h->err, st, count, true));
    } else {
        gps::ScanParams p{};
        const uint64_t tile = h->shape.tile();
        p.in = d_in;
        p.out = d_out;
        p.n = n;
        p.num_tiles = (uint32_t)((n + tile - 1) / tile);
        p.vec_ok = aligned16(d_in) && aligned16(d_out);
        p.status = h->status;
        p.ticket = h->ticket;
        p.err = h->err;
        p.occupancy = count;
        *resident = gps::pp_resident_workgroups(h->shape, true);
        HIP_TRY(gps::launch_csdl_pp_u32(h->shape, true, p, st, true));
    }
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, count, sizeof c, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *measured = (uint32_t)c;
    return GPS_OK;
}

int gps_inclusive_scan_u32x4(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream) {
    return scan_u32x4(h, d_in, d_out, n, true, stream);
}

int gps_exclusive_scan_u32x4(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream) {
    return scan_u32x4(h, d_in, d_out, n, false, stream);
}

int gps_inclusive_scan_fp_per_segment(gps_handle* h, const void* d_in, float* d_out, const uint64_t* h_offsets,
                                      uint64_t num_segments, int bf16, int compensated, void* stream) {
    if (!h || !h_offsets) return fail(GPS_ERR_INVALID, "null argument");
    const size_t esize = bf16 ? 2 : 4;
    for (uint64_t s = 0; s < num_segments; ++s) {
        const uint64_t a = h_offsets[s], b = h_offsets[s + 1];
        if (b < a) return fail(GPS_ERR_INVALID, "offsets must be non-decreasing (segment %llu)", (unsigned long long)s);
        if (b == a) continue;
        const void* in = static_cast<const char*>(d_in) + a * esize;
        const int rc = scan_fp(h, in, d_out + a, b - a, bf16 != 0, compensated, stream);
        if (rc) return rc;
    }
    return GPS_OK;
}

int gps_graph_launch(gps_graph* g, void* stream) {
    if (!g || !g->exec) return fail(GPS_ERR_INVALID, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    return GPS_OK;
}

int gps_graph_destroy(gps_graph* g) {
    if (!g) return fail(GPS_ERR_INVALID, "null graph");
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return GPS_OK;
}

}  // extern "C"
