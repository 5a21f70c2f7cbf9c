// The C-ABI handle and the error helpers shared by the entry-point files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gps.h"
#include "gps_internal.hpp"

// Lower bound on the tile of every kernel shape (the smallest is the
// control-wave schedule's 3 data waves x 2 uint4 = 1536 elements; the FP
// scans use two words per tile of at least 4096): status words are sized for
// it so every entry point of a handle, whatever shape it runs, has its words.
constexpr uint32_t kMinTileElems = 1024;

namespace gps {

// A handle's ticket word (first 256 bytes) and status slots (the rest), in
// one block of device memory; see status_memory.cpp.
struct StatusBlock {
    void* ptr = nullptr;
    uint64_t bytes = 0;
    bool uncached = false;
};
constexpr uint64_t kTicketBytes = 256;

// Takes a block of at least `bytes` for `device` (the current device):
// uncached memory from the per-device free list or newly allocated, or
// ordinary memory when GPS_STATUS_MEMORY=plain.
hipError_t status_block_acquire(int device, uint64_t bytes, StatusBlock& out);
// Returns an uncached block to the free list (never to the runtime); frees an
// ordinary one.
void status_block_release(int device, const StatusBlock& b);
// Bytes of uncached blocks allocated so far, and of those on the free lists.
void status_memory_stats(uint64_t* held, uint64_t* free_bytes);

}  // namespace gps

struct gps_handle {
    int algo = 0;
    uint64_t max_n = 0;
    gps::TileShape shape{256, 4};
    uint32_t spin_limit = 4;
    uint32_t stall_period = 0;
    uint32_t schedule = GPS_SCHEDULE_TILE;
    bool size_adaptive = false;  // CSDL with default shape and schedule: both picked per call from n
    uint32_t max_tiles = 0;     // status slots allocated (one per kMinTileElems elements)
    int device = 0;
    gps::StatusBlock status_block;  // holds ticket and status
    uint64_t* status = nullptr;
    uint64_t* ticket = nullptr;
    uint32_t* err = nullptr;
    unsigned long long* counters = nullptr;
    uint32_t* tile_sums = nullptr;  // reduce-then-scan spine
    uint64_t launches = 0;          // epoch-tagged launches since the last re-arm
};

namespace gps {

// Records the message for gps_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
// GPS_OK when the handle's device is the current one, else a GPS_ERR_INVALID
// failure (its buffers live on that device; a launch elsewhere would fault).
int check_device(const gps_handle* h);
// Re-arms the status words when the 30-bit epoch would wrap; counts a launch.
hipError_t before_epoch_launch(gps_handle* h, hipStream_t s);

}  // namespace gps

#define HIP_TRY(expr)                                          \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return gps::hip_fail(e_, #expr); \
    } while (0)
