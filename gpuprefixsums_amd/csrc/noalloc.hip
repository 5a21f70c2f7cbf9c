// Allocation-free uint32 scan (gps_scan_u32_noalloc in include/gps.h).
//
// The handle-based entry points own their scratch (status slots, ticket and
// error words) in device memory that gps_create allocates. Callers that
// cannot allocate (code under stream capture, serving loops, harnesses that
// forbid hipMalloc) use this entry point instead: its scratch is the
// file-scope device array `scratch` below, of a fixed size, one copy per GPU
// for the life of the process. The entry point calls no allocating routine.
//
// Setup, once per device on the first call made with that device current:
// hipGetSymbolAddress gives this device's address of `scratch` (the array's
// name in host code is a host-side stand-in, not a device address, so it is
// never passed to a kernel or a copy), and the status slots and ticket word
// are re-armed on the caller's stream. After that every call uses the
// recorded address.
//
// The slots need no clearing between calls: every status word carries the
// epoch of the launch that wrote it and the ticket word re-arms itself
// (device_common.hpp), so each call is the plain scan's one kernel launch.
// Every call on a GPU shares that GPU's array: calls that overlap in time on
// one GPU (several streams or host threads) race and must be serialised by
// the caller.
#include <atomic>
#include <mutex>

#include "handle.hpp"

namespace {

// Largest n: 2^28 elements (1 GiB of uint32). The default configuration
// for n <= 2^28 runs tiles of at least 8192 elements and at most 8192 tiles
// (persistent 1024 x 8 tiles of 32768 elements from 2^24 up), so 8192
// status slots cover every size up to the limit.
constexpr uint64_t kMaxN = 1ull << 28;
constexpr uint64_t kSlots = 8192;
constexpr uint64_t kStatusWords = kSlots * gps::kSlotWords;
constexpr uint64_t kTicketWord = kStatusWords;      // (epoch << 32) | tickets taken
constexpr uint64_t kErrWord = kStatusWords + 16;    // watchdog bits (own 128-byte line)
constexpr uint64_t kScratchWords = kStatusWords + 32;

// 1 MiB + 256 B per GPU.
static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::once_flag g_setup_once[kMaxDevices];
uint64_t* g_scratch[kMaxDevices];             // this device's address of `scratch`
hipError_t g_setup_error[kMaxDevices];
std::atomic<uint64_t> g_launches[kMaxDevices];  // epoch-tagged launches since the last re-arm
constexpr uint64_t kLaunchesBeforeReset = (1ull << 30) - 2;

// Once per device, with that device current: look up the device address of
// `scratch` and re-arm its status slots and ticket word on `s`.
void setup(int dev, hipStream_t s) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
    if (e == hipSuccess) {
        uint64_t* w = static_cast<uint64_t*>(p);
        e = gps::launch_reset_status(w, kStatusWords, w + kTicketWord, s);
        if (e == hipSuccess) e = gps::launch_fill_u32(reinterpret_cast<uint32_t*>(w + kErrWord), 2, 0u, s);
        if (e == hipSuccess) g_scratch[dev] = w;
    }
    g_setup_error[dev] = e;
}

}  // namespace

using gps::fail;
using gps::hip_fail;

extern "C" int gps_scan_u32_noalloc(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* stream) {
    if (n > kMaxN)
        return fail(GPS_ERR_TOO_LARGE, "n=%llu exceeds the allocation-free scan's limit of %llu elements",
                    (unsigned long long)n, (unsigned long long)kMaxN);
    if (n && (!d_in || !d_out)) return fail(GPS_ERR_INVALID, "null data pointer with n=%llu", (unsigned long long)n);
    if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & 3u)
        return fail(GPS_ERR_INVALID, "uint32 buffers must be 4-byte aligned");
    if (n == 0) return GPS_OK;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GPS_ERR_INVALID, "device %d out of range", dev);
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::call_once(g_setup_once[dev], setup, dev, s);
    if (g_setup_error[dev] != hipSuccess) return hip_fail(g_setup_error[dev], "allocation-free scan setup");
    uint64_t* w = g_scratch[dev];

    if (g_launches[dev].fetch_add(1) >= kLaunchesBeforeReset) {  // the 30-bit epoch would wrap
        HIP_TRY(gps::launch_reset_status(w, kStatusWords, w + kTicketWord, s));
        g_launches[dev].store(1);
    }
    gps::TileShape shape{1024, 8};
    uint32_t schedule = GPS_SCHEDULE_PERSISTENT;
    gps::size_adaptive_config(n, shape, schedule);
    const uint64_t tiles = (n + shape.tile() - 1) / shape.tile();
    if (tiles > kSlots) return fail(GPS_ERR_TOO_LARGE, "%llu tiles exceed the %llu status slots",
                                    (unsigned long long)tiles, (unsigned long long)kSlots);
    gps::ScanParams p{};
    p.in = d_in;
    p.out = d_out;
    p.n = n;
    p.num_tiles = (uint32_t)tiles;
    p.vec_ok = ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & 15u) == 0;
    p.status = w;
    p.ticket = w + kTicketWord;
    p.err = reinterpret_cast<uint32_t*>(w + kErrWord);
    if (schedule == GPS_SCHEDULE_PERSISTENT)
        HIP_TRY(gps::launch_csdl_pp_u32(shape, inclusive != 0, p, s));
    else
        HIP_TRY(gps::launch_csdl_u32(shape, inclusive != 0, p, s));
    return GPS_OK;
}
