// C ABI of libgps.so (declared in include/gps.h): handles, argument checks,
// error reporting and dispatch to the kernel launchers.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "handle.hpp"

namespace {

thread_local std::string g_last_error;

// Epochs are 30 bits wide; re-arm the status words before a value could recur.
constexpr uint64_t kLaunchesBeforeReset = (1ull << 30) - 2;

}  // namespace

namespace gps {

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(GPS_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int check_device(const gps_handle* h) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != h->device)
        return fail(GPS_ERR_INVALID, "the handle belongs to device %d but device %d is current", h->device, cur);
    return GPS_OK;
}

hipError_t before_epoch_launch(gps_handle* h, hipStream_t s) {
    if (h->launches >= kLaunchesBeforeReset) {
        const hipError_t e = launch_reset_status(h->status, (uint64_t)h->max_tiles * kSlotWords, h->ticket, s);
        if (e != hipSuccess) return e;
        h->launches = 0;
    }
    ++h->launches;
    return hipSuccess;
}

}  // namespace gps

using gps::fail;
using gps::hip_fail;

namespace {

// Defaults from tile-shape sweeps at n = 2^28 on MI355X. Single-pass scans
// want large tiles: each tile costs one device-scope ticket atomic and one
// status round trip, both serialised at the coherence point.
gps::TileShape default_shape(int algo) {
    switch (algo) {
        case GPS_ALGO_RTS:
            return {512, 2};
        case GPS_ALGO_CSDLDF:
            return {1024, 8};  // 5767 / 4218 / 4248 GB/s at stall periods 0 / 7 / 8 (1024 x 4: 5093 / 3293 / 3304)
        default:
            return {1024, 8};
    }
}

// Elements per tile of the plain-scan kernel for a shape and schedule.
uint32_t kernel_tile(gps::TileShape shape, uint32_t schedule) {
    return schedule == GPS_SCHEDULE_CONTROL ? gps::cw_tile_elems(shape) : shape.tile();
}

uint32_t default_schedule(int algo) { return algo == GPS_ALGO_CSDL ? GPS_SCHEDULE_PERSISTENT : GPS_SCHEDULE_TILE; }

void free_handle(gps_handle* h) {
    if (!h) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(h->device);
    // hipFree waits for the device; an uncached status block goes back to a
    // free list instead, so wait here: no launch still in flight may touch
    // the block once another handle can take it.
    (void)hipDeviceSynchronize();
    gps::status_block_release(h->device, h->status_block);
    (void)hipFree(h->err);
    (void)hipFree(h->counters);
    (void)hipFree(h->tile_sums);
    (void)hipSetDevice(cur);
    delete h;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

int check_scan_args(gps_handle* h, const void* in, const void* out, uint64_t n) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    if (n && (!in || !out)) return fail(GPS_ERR_INVALID, "null data pointer with n=%llu", (unsigned long long)n);
    if (n > h->max_n)
        return fail(GPS_ERR_TOO_LARGE, "n=%llu exceeds handle max_n=%llu", (unsigned long long)n,
                    (unsigned long long)h->max_n);
    if ((reinterpret_cast<uintptr_t>(in) & 3u) || (reinterpret_cast<uintptr_t>(out) & 3u))
        return fail(GPS_ERR_INVALID, "uint32 buffers must be 4-byte aligned");
    return GPS_OK;
}

int scan_u32(gps_handle* h, const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive, void* stream,
             const uint32_t* carry_in = nullptr, uint32_t* carry_out = nullptr) {
    int rc = check_scan_args(h, in, out, n);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {
        if (carry_out) {
            if (carry_in)
                HIP_TRY(hipMemcpyAsync(carry_out, carry_in, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
            else  // a fill kernel, not a memset: the call stays capturable into a graph
                HIP_TRY(gps::launch_fill_u32(carry_out, 1, 0u, s));
        }
        return GPS_OK;
    }
    gps::TileShape shape = h->shape;
    uint32_t schedule = h->schedule;
    if (h->size_adaptive) gps::size_adaptive_config(n, shape, schedule);
    if ((carry_in || carry_out) &&
        (h->algo != GPS_ALGO_CSDL || schedule == GPS_SCHEDULE_CONTROL))
        return fail(GPS_ERR_UNSUPPORTED, "carries are offered by the CSDL tile and persistent schedules");
    const uint32_t tile = kernel_tile(shape, schedule);
    const uint32_t tiles = (uint32_t)((n + tile - 1) / tile);
    const uint32_t vec_ok = aligned16(in) && aligned16(out);
    if (h->algo == GPS_ALGO_RTS) {
        HIP_TRY(gps::launch_rts_u32(shape, inclusive, in, out, n, tiles, vec_ok, h->tile_sums, s));
        return GPS_OK;
    }
    HIP_TRY(gps::before_epoch_launch(h, s));
    gps::ScanParams p{};
    p.in = in;
    p.out = out;
    p.n = n;
    p.num_tiles = tiles;
    p.vec_ok = vec_ok;
    p.status = h->status;
    p.ticket = h->ticket;
    p.err = h->err;
    p.spin_limit = h->spin_limit;
    p.stall_period = h->stall_period;
    p.counters = h->counters;
    p.carry_in = carry_in;
    p.carry_out = carry_out;
    if (h->algo == GPS_ALGO_CSDLDF) {
        if (schedule == GPS_SCHEDULE_PERSISTENT) {
            HIP_TRY(gps::launch_csdldf_pp_u32(shape, inclusive, p, s));
        } else if (schedule == GPS_SCHEDULE_SPLIT) {
            // Both launches share one epoch: the second is counted here too,
            // so no host re-arm can fall between them.
            HIP_TRY(gps::before_epoch_launch(h, s));
            HIP_TRY(gps::launch_csdldf_split_u32(shape, inclusive, p, s));
        } else
            HIP_TRY(gps::launch_csdldf_u32(shape, inclusive, p, s));
    } else if (schedule == GPS_SCHEDULE_PERSISTENT) {
        HIP_TRY(gps::launch_csdl_pp_u32(shape, inclusive, p, s));
    } else if (schedule == GPS_SCHEDULE_CONTROL) {
        HIP_TRY(gps::launch_csdl_cw_u32(shape, inclusive, p, s));
    } else {
        HIP_TRY(gps::launch_csdl_u32(shape, inclusive, p, s));
    }
    return GPS_OK;
}

}  // namespace

extern "C" {

const char* gps_version(void) { return "gpuprefixsums_amd 0.1 (gfx950)"; }

const char* gps_last_error(void) { return g_last_error.c_str(); }

int gps_create(int algo, uint64_t max_n, const gps_tuning* tuning, gps_handle** out) {
    if (!out) return fail(GPS_ERR_INVALID, "null output handle pointer");
    *out = nullptr;
    if (algo != GPS_ALGO_CSDL && algo != GPS_ALGO_RTS && algo != GPS_ALGO_CSDLDF)
        return fail(GPS_ERR_INVALID, "unknown algorithm %d", algo);
    if (max_n == 0) return fail(GPS_ERR_INVALID, "max_n must be positive");
    gps::TileShape shape = default_shape(algo);
    uint32_t spin_limit = 4, stall_period = 0, schedule = GPS_SCHEDULE_DEFAULT;
    if (tuning) {
        if (tuning->block_threads) shape.block = tuning->block_threads;
        if (tuning->vec_per_thread) shape.vec = tuning->vec_per_thread;
        if (tuning->spin_limit) spin_limit = tuning->spin_limit;
        stall_period = tuning->stall_period;
        schedule = tuning->schedule;
    }
    if (schedule > GPS_SCHEDULE_SPLIT) return fail(GPS_ERR_INVALID, "unknown schedule %u", schedule);
    if (schedule == GPS_SCHEDULE_SPLIT && (algo != GPS_ALGO_CSDLDF || stall_period < 2))
        return fail(GPS_ERR_INVALID, "the split schedule needs the decoupled-fallback algorithm and stall_period >= 2");
    if (schedule == GPS_SCHEDULE_PERSISTENT && algo == GPS_ALGO_RTS)
        return fail(GPS_ERR_INVALID, "the persistent schedule is offered for CSDL and CSDLDF");
    if (schedule == GPS_SCHEDULE_CONTROL && algo != GPS_ALGO_CSDL)
        return fail(GPS_ERR_INVALID, "the control-wave schedule is offered for CSDL only");
    const bool size_adaptive = algo == GPS_ALGO_CSDL && schedule == GPS_SCHEDULE_DEFAULT &&
                               !(tuning && (tuning->block_threads || tuning->vec_per_thread));
    if (schedule == GPS_SCHEDULE_DEFAULT) schedule = default_schedule(algo);
    if (!gps::shape_supported(shape))
        return fail(GPS_ERR_INVALID, "unsupported tile shape block=%u vec=%u (block 256/512/1024, vec 2/4/8)",
                    shape.block, shape.vec);
    if (stall_period == 1) return fail(GPS_ERR_INVALID, "stall_period must be 0 or >= 2");
    // A waiter polls up to 8 x spin_limit times; the bound keeps that product
    // in 32 bits and below the device watchdog's 2^20 polls.
    if (spin_limit > 65536) return fail(GPS_ERR_INVALID, "spin_limit must be at most 65536");
    if (stall_period && algo != GPS_ALGO_CSDLDF)
        return fail(GPS_ERR_INVALID, "stall emulation needs the decoupled-fallback algorithm");
    // One status slot (kSlotWords words) per kMinTileElems elements: every
    // kernel's tile holds at least that many elements (structs for the struct
    // scan), so a launch over n <= max_n never needs more slots.
    const uint64_t tiles = (max_n + kMinTileElems - 1) / kMinTileElems;
    if (tiles >= (1ull << 31)) return fail(GPS_ERR_TOO_LARGE, "max_n too large");

    gps_handle* h = new gps_handle();
    h->algo = algo;
    h->max_n = max_n;
    h->shape = shape;
    h->spin_limit = spin_limit;
    h->stall_period = stall_period;
    h->schedule = schedule;
    h->size_adaptive = size_adaptive;
    h->max_tiles = (uint32_t)tiles;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess)
        e = gps::status_block_acquire(h->device, gps::kTicketBytes + tiles * gps::kSlotWords * sizeof(uint64_t),
                                      h->status_block);
    if (e == hipSuccess) {
        h->ticket = static_cast<uint64_t*>(h->status_block.ptr);
        h->status = reinterpret_cast<uint64_t*>(static_cast<char*>(h->status_block.ptr) + gps::kTicketBytes);
    }
    if (e == hipSuccess) e = hipMalloc(&h->err, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&h->counters, gps::kCounterWords * sizeof(unsigned long long));
    if (e == hipSuccess && algo == GPS_ALGO_RTS) e = hipMalloc(&h->tile_sums, tiles * sizeof(uint32_t));
    if (e == hipSuccess) e = gps::launch_reset_status(h->status, tiles * gps::kSlotWords, h->ticket, nullptr);
    if (e == hipSuccess) e = hipMemset(h->err, 0, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(h->counters, 0, gps::kCounterWords * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_handle(h);
        return hip_fail(e, "gps_create");
    }
    *out = h;
    return GPS_OK;
}

int gps_destroy(gps_handle* h) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    free_handle(h);
    return GPS_OK;
}

int gps_get_tuning(const gps_handle* h, gps_tuning* out, uint32_t* tile_elems) {
    if (!h || !out) return fail(GPS_ERR_INVALID, "null argument");
    out->block_threads = h->shape.block;
    out->vec_per_thread = h->shape.vec;
    out->spin_limit = h->spin_limit;
    out->stall_period = h->stall_period;
    out->schedule = h->schedule;
    if (tile_elems) *tile_elems = kernel_tile(h->shape, h->schedule);
    return GPS_OK;
}

int gps_inclusive_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream) {
    return scan_u32(h, d_in, d_out, n, true, stream);
}

int gps_exclusive_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream) {
    return scan_u32(h, d_in, d_out, n, false, stream);
}

int gps_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* stream) {
    return scan_u32(h, d_in, d_out, n, inclusive != 0, stream);
}

int gps_scan_u32_carry(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive,
                       const uint32_t* d_carry_in, uint32_t* d_carry_out, void* stream) {
    if ((reinterpret_cast<uintptr_t>(d_carry_in) | reinterpret_cast<uintptr_t>(d_carry_out)) & 3u)
        return fail(GPS_ERR_INVALID, "carry words must be 4-byte aligned");
    return scan_u32(h, d_in, d_out, n, inclusive != 0, stream, d_carry_in, d_carry_out);
}

int gps_get_counters(gps_handle* h, uint64_t out[4]) {
    if (!h || !out) return fail(GPS_ERR_INVALID, "null argument");
    unsigned long long c[4];
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(c, h->counters, sizeof c, hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; ++i) out[i] = c[i];
    return GPS_OK;
}

int gps_reset_counters(gps_handle* h) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemset(h->counters, 0, 4 * sizeof(unsigned long long)));
    HIP_TRY(hipDeviceSynchronize());
    return GPS_OK;
}

int gps_device_error(gps_handle* h, uint32_t* bits) {
    if (!h || !bits) return fail(GPS_ERR_INVALID, "null argument");
    if (int rc_ = gps::check_device(h)) return rc_;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(bits, h->err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(h->err, 0, sizeof(uint32_t)));
    if (*bits) {
        // A launch that raised an error may have left tiles unpublished or its
        // ticket word short of the count that re-arms it: start the next
        // launch from cleared status words and a fresh ticket word.
        HIP_TRY(gps::launch_reset_status(h->status, (uint64_t)h->max_tiles * gps::kSlotWords, h->ticket, nullptr));
        h->launches = 0;
    }
    HIP_TRY(hipDeviceSynchronize());
    return GPS_OK;
}

int gps_reset(gps_handle* h, void* stream) {
    if (!h) return fail(GPS_ERR_INVALID, "null handle");
    if (int rc_ = gps::check_device(h)) return rc_;
    HIP_TRY(gps::launch_reset_status(h->status, (uint64_t)h->max_tiles * gps::kSlotWords, h->ticket,
                                     static_cast<hipStream_t>(stream)));
    h->launches = 0;
    return GPS_OK;
}

int gps_time_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive,
                      void* stream, uint32_t iters, float* ms) {
    if (!h || !ms || iters == 0) return fail(GPS_ERR_INVALID, "bad timing arguments");
    int rc = check_scan_args(h, d_in, d_out, n);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<hipEvent_t> ev(iters + 1);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    int status = GPS_OK;
    HIP_TRY(hipEventRecord(ev[0], s));
    for (uint32_t i = 0; i < iters && status == GPS_OK; ++i) {
        status = scan_u32(h, d_in, d_out, n, inclusive != 0, stream);
        if (status == GPS_OK) {
            hipError_t e = hipEventRecord(ev[i + 1], s);
            if (e != hipSuccess) status = hip_fail(e, "hipEventRecord");
        }
    }
    if (status == GPS_OK) {
        hipError_t e = hipEventSynchronize(ev[iters]);
        if (e != hipSuccess) status = hip_fail(e, "hipEventSynchronize");
        for (uint32_t i = 0; i < iters && status == GPS_OK; ++i) {
            e = hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
            if (e != hipSuccess) status = hip_fail(e, "hipEventElapsedTime");
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return status;
}

int gps_fill_u32(uint32_t* d_ptr, uint64_t n, uint32_t value, void* stream) {
    if (n && !d_ptr) return fail(GPS_ERR_INVALID, "null pointer");
    HIP_TRY(gps::launch_fill_u32(d_ptr, n, value, static_cast<hipStream_t>(stream)));
    return GPS_OK;
}

int gps_wave_primitive_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int op, void* stream) {
    if (n && (!d_in || !d_out)) return fail(GPS_ERR_INVALID, "null pointer");
    if (op < GPS_WAVE_INCLUSIVE_SCAN || op > GPS_WAVE_EXCLUSIVE_SCAN) return fail(GPS_ERR_INVALID, "unknown op %d", op);
    if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & 3u)
        return fail(GPS_ERR_INVALID, "uint32 buffers must be 4-byte aligned");
    HIP_TRY(gps::launch_wave_unit_u32(d_in, d_out, n, (uint32_t)op, static_cast<hipStream_t>(stream)));
    return GPS_OK;
}

int gps_copy_u32(uint32_t* d_dst, const uint32_t* d_src, uint64_t n, void* stream) {
    if (n && (!d_dst || !d_src)) return fail(GPS_ERR_INVALID, "null pointer");
    if ((reinterpret_cast<uintptr_t>(d_dst) | reinterpret_cast<uintptr_t>(d_src)) & 3u)
        return fail(GPS_ERR_INVALID, "uint32 buffers must be 4-byte aligned");
    HIP_TRY(gps::launch_copy_u32(d_src, d_dst, n, static_cast<hipStream_t>(stream)));
    return GPS_OK;
}

int gps_fill_random_u32(uint32_t* d_ptr, uint64_t n, uint64_t seed, uint64_t first, uint32_t bits,
                        void* stream) {
    if (n && !d_ptr) return fail(GPS_ERR_INVALID, "null pointer");
    if (bits < 1 || bits > 32) return fail(GPS_ERR_INVALID, "bits must be in [1, 32]");
    HIP_TRY(gps::launch_fill_splitmix_u32(d_ptr, n, seed, first, bits, static_cast<hipStream_t>(stream)));
    return GPS_OK;
}

static int run_validator(bool ones, const uint32_t* in, const uint32_t* scan, uint64_t n, int inclusive,
                         void* stream, uint64_t* errors) {
    if (!errors || (n && (!scan || (!ones && !in)))) return fail(GPS_ERR_INVALID, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* d_err = nullptr;
    HIP_TRY(hipMalloc(&d_err, sizeof *d_err));
    hipError_t e = hipMemsetAsync(d_err, 0, sizeof *d_err, s);
    if (e == hipSuccess)
        e = ones ? gps::launch_validate_ones_u32(scan, n, inclusive, d_err, s)
                 : gps::launch_validate_diff_u32(in, scan, n, inclusive, d_err, s);
    unsigned long long h_err = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h_err, d_err, sizeof h_err, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_err);
    if (e != hipSuccess) return hip_fail(e, "validator");
    *errors = h_err;
    return GPS_OK;
}

int gps_validate_ones_u32(const uint32_t* d_scan, uint64_t n, int inclusive, void* stream, uint64_t* errors) {
    return run_validator(true, nullptr, d_scan, n, inclusive, stream, errors);
}

int gps_validate_scan_u32(const uint32_t* d_in, const uint32_t* d_scan, uint64_t n, int inclusive,
                          void* stream, uint64_t* errors) {
    return run_validator(false, d_in, d_scan, n, inclusive, stream, errors);
}

}  // extern "C"
