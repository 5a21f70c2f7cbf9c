// Internal (host-side) interface between the C ABI layer and the kernel
// launchers. Not installed; the public surface is include/gps.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gps {

// Tile status words are laid out one tile per 128-byte slot of kSlotWords
// words (a tile with several words keeps them all in its slot), so the
// publications of neighbouring tiles, made from different XCDs, never share a
// cache line. A handle holds one slot per kMinTileElems elements of max_n.
constexpr uint64_t kSlotWords = 16;

// Handle counter words: [0..3] the decoupled-fallback statistics, [4] the
// occupancy estimate's count of workgroups that took a tile.
constexpr uint64_t kCounterWords = 8;
constexpr uint64_t kCounterOccupancy = 4;

// Launch parameters shared by the single-pass uint32 kernels (CSDL, CSDLDF).
struct ScanParams {
    const uint32_t* in;
    uint32_t* out;
    uint64_t n;
    uint32_t num_tiles;
    uint32_t vec_ok;  // in and out are 16-byte aligned
    uint64_t* status;  // one word per tile
    uint64_t* ticket;  // (epoch << 32) | tickets taken
    uint32_t* err;     // device error bits (watchdogs)
    // decoupled fallback only
    uint32_t spin_limit;
    uint32_t stall_period;              // 0: no emulated stalls
    unsigned long long* counters;       // [polls, fallbacks, insertions, stall timeouts] or null
    // persistent CSDL only: +1 per workgroup that takes a tile, or null
    unsigned long long* occupancy;
    // CSDL (tile and persistent schedules): a device word added to every
    // output (null: 0), and a device word that receives the scan's total plus
    // that carry (null: not written). The carry is read before tile 0
    // publishes and the total written after the last tile's lookback, so the
    // two may be the same word (chunked scans chain through it).
    const uint32_t* carry_in;
    uint32_t* carry_out;
    // CSDLDF split schedule (two launches, the D3D12 deadlock emulation):
    // 0 = not split, 1 = first launch (tiles with index % stall_period != 0),
    // 2 = second launch (the deferred tiles, under the first launch's epoch)
    uint32_t split_pass;
};

// Workgroups of KERNEL (BLOCK threads) resident on the current device at
// once: the grid of a persistent launch. Queried once per kernel and process.
template <auto KERNEL, int BLOCK>
uint32_t resident_workgroups() {
    struct Residency {
        uint32_t workgroups = 0;
        Residency() {
            int dev = 0, cus = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess) return;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, KERNEL, BLOCK, 0) != hipSuccess) return;
            workgroups = (uint32_t)(cus > 0 ? cus : 1) * (uint32_t)(per_cu > 0 ? per_cu : 1);
        }
    };
    static const Residency r;
    return r.workgroups;
}

// Grid of a persistent launch over `tiles` tiles with `resident` workgroups
// resident: every resident workgroup, or one per tile when there are fewer
// tiles. (Trimming the grid so the last round of tiles is nearly full, the
// fewest workgroups that need ceil(tiles / resident) rounds, measured within
// 1.5% at 50.9 M elements and slower for the segmented kernels.)
inline uint32_t persistent_grid(uint32_t tiles, uint32_t resident) { return tiles < resident ? tiles : resident; }

// Compute units of the current device (queried once per process; 0 if the
// query fails).
inline uint32_t device_cus() {
    static const uint32_t cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0u;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0u;
        return (uint32_t)(n > 0 ? n : 0);
    }();
    return cus;
}

struct TileShape {
    uint32_t block;
    uint32_t vec;
    uint32_t tile() const { return block * vec * 4u; }
};

bool shape_supported(TileShape s);

// Shape and schedule of a default CSDL handle for an n-element scan (the
// schedule constants of gps.h: 1 = one workgroup per tile, 2 = persistent),
// measured on MI355X (back-to-back launches): below 2^24 elements one
// workgroup per tile beats the persistent grid, which has too few tiles per
// workgroup to pay for its ramp; below 2^22 smaller tiles also shorten the
// lookback chain. 2^16: 3.6 against 5.5 us; 2^20: 6.1 against 7.4 us; 2^23:
// 14.4 against 15.4 us; from 2^24 up the persistent 1024 x 8 kernel wins
// (26.7 against 28.9 us at 2^24). Leaves shape and schedule unchanged from
// 2^24 up (the persistent 1024 x 8 default).
inline void size_adaptive_config(uint64_t n, TileShape& shape, uint32_t& schedule) {
    constexpr uint32_t kTile = 1;  // GPS_SCHEDULE_TILE
    if (n <= (1ull << 18)) {
        shape = {1024, 2};
        schedule = kTile;
    } else if (n <= (1ull << 21)) {
        shape = {1024, 4};
        schedule = kTile;
    } else if (n < (1ull << 24)) {
        shape = {1024, 8};
        schedule = kTile;
    }
}

// Each launcher enqueues exactly one kernel (RTS: three) on `stream` and
// returns the HIP launch status.
hipError_t launch_csdl_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream);
// Persistent pipelined CSDL: grid = resident workgroups (<= num_tiles).
// With grid_all_tiles, one workgroup is launched per tile instead: those that
// start after the tickets run out exit at once (the occupancy estimate).
hipError_t launch_csdl_pp_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream,
                              bool grid_all_tiles = false);
// Workgroups the persistent kernel keeps resident (the occupancy query times
// the CU count); 0 for an unsupported shape or a failed query.
uint32_t pp_resident_workgroups(TileShape s, bool inclusive);
// Persistent CSDL with a control wave: block - 64 data threads per tile.
uint32_t cw_tile_elems(TileShape s);
hipError_t launch_csdl_cw_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream);
hipError_t launch_csdldf_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream);
// Split schedule: the first launch (every tile except index % stall_period
// == 0) and then the second (the deferred tiles), back to back on `stream`.
hipError_t launch_csdldf_split_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream);
// Persistent decoupled-fallback scan (grid = resident workgroups).
hipError_t launch_csdldf_pp_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream);
hipError_t launch_rts_u32(TileShape s, bool inclusive, const uint32_t* in, uint32_t* out, uint64_t n,
                          uint32_t num_tiles, uint32_t vec_ok, uint32_t* tile_sums, hipStream_t stream);

// Batched segmented scan in one launch: persistent (grid = resident
// workgroups, pipelined) or one workgroup per tile.
bool seg_shape_supported(TileShape s, bool persistent);
// Runs the persistent kernel's one-time occupancy query (call before a
// stream capture, so the capture holds only the launch).
hipError_t prepare_segscan_u32(TileShape s, bool inclusive, bool persistent);
hipError_t launch_segscan_u32(TileShape s, bool inclusive, bool persistent, const uint32_t* in, uint32_t* out,
                              const uint64_t* offsets, uint64_t num_segments, uint64_t n, uint32_t vec_ok,
                              uint64_t* status, uint64_t* ticket, uint32_t* err, hipStream_t stream);

// FP32 / BF16-input inclusive scans with FP32 output (2 status words per tile).
bool fp_shape_supported(TileShape s);
// Carries of the FP scans (values match GPS_FP_* in gps.h).
enum FpCarry : int { kFpPlain = 0, kFpPair = 1 };
// persistent: resident workgroups loop over ticketed tiles, loading tile i+1
// while tile i's lookback resolves (otherwise one workgroup per tile).
hipError_t launch_scan_fp(TileShape s, bool bf16, int carry, bool persistent, const void* in, float* out,
                          uint64_t n, uint32_t vec_ok, uint64_t* status, uint64_t* ticket, uint32_t* err,
                          hipStream_t stream);
// Batched segmented FP inclusive scan (one launch; same shapes as launch_scan_fp;
// persistent as there).
hipError_t launch_segscan_fp(TileShape s, bool bf16, bool compensated, bool persistent, const void* in, float* out,
                             const uint64_t* offsets, uint64_t num_segments, uint64_t n, uint32_t vec_ok,
                             uint64_t* status, uint64_t* ticket, uint32_t* err, hipStream_t stream);

// Member-wise scan of n 4-member uint32 structs (4 status words per tile).
bool v4_shape_supported(TileShape s);
// persistent: resident workgroups with the pipelined ticket loop; otherwise
// one workgroup per tile.
hipError_t launch_csdl_u32x4(TileShape s, bool inclusive, bool persistent, const uint32_t* in, uint32_t* out,
                             uint64_t n, uint64_t* status, uint64_t* ticket, uint32_t* err, hipStream_t stream,
                             unsigned long long* occupancy = nullptr, bool grid_all_tiles = false);
uint32_t v4_resident_workgroups(TileShape s, bool inclusive);

// Survey of wave and block scan networks over fixed-length segments
// (values match gps_survey_alg in gps.h).
enum SurveyAlg : int {
    kSurveySerial = 0,
    kSurveyWaveKoggeStoneLds = 1,
    kSurveyWaveKoggeStoneShuffle = 2,
    kSurveyWaveDpp = 3,
    kSurveyWaveRaking = 4,
    kSurveyBlockKoggeStone = 5,
    kSurveyBlockSklansky = 6,
    kSurveyBlockBrentKung = 7,
    kSurveyBlockWaveRaking = 8,
    kSurveyBlockWaveSklansky = 9,
    kSurveyTileRegisters = 10,
    kSurveyCount = 11,
};
// Segment length in elements (0 for an unknown network).
uint32_t survey_segment(int alg);
hipError_t launch_survey_u32(int alg, bool inclusive, const uint32_t* in, uint32_t* out, uint64_t n,
                             hipStream_t stream);

// Clears `count` status words and arms the ticket word at epoch 1.
hipError_t launch_reset_status(uint64_t* status, uint64_t count, uint64_t* ticket, hipStream_t stream);

// Utility kernels (the reference's InitOne / Validate* family and a
// counter-based random fill).
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t value, hipStream_t stream);
hipError_t launch_copy_u32(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t stream);
hipError_t launch_fill_splitmix_u32(uint32_t* p, uint64_t n, uint64_t seed, uint64_t first, uint32_t bits,
                                    hipStream_t stream);
// Wave64 primitives over runs of 64 consecutive elements (device unit test).
hipError_t launch_wave_unit_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t op, hipStream_t stream);
hipError_t launch_validate_ones_u32(const uint32_t* scan, uint64_t n, int inclusive,
                                    unsigned long long* errors, hipStream_t stream);
hipError_t launch_validate_diff_u32(const uint32_t* in, const uint32_t* scan, uint64_t n, int inclusive,
                                    unsigned long long* errors, hipStream_t stream);

}  // namespace gps
