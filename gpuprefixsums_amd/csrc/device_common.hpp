// Device building blocks shared by the gfx950 scan kernels: wave64 DPP scans,
// 64-bit tile status words with per-launch epochs, and ticketed tile order.
//
// Reference counterparts (behaviour, not code):
//   wave scans      <- GPUPrefixSumsCUDA/Utils.cuh:29-59 (warp-32 shuffle scans)
//   status flags    <- GPUPrefixSumsCUDA/ChainedScanDecoupledLookback.cuh:19-23
//   ticketed tiles  <- GPUPrefixSumsCUDA/ChainedScanDecoupledLookback.cuh:89-94
//   window lookback <- GPUPrefixSumsCUDA/ChainedScanDecoupledLookback.cuh:50-77
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gps {

constexpr uint32_t kWave = 64;

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Wave index inside the workgroup, in a scalar register: branches on it are
// uniform to the compiler, so role-split code (one wave polls, the others
// stream) is not structured as a divergent region.
__device__ __forceinline__ uint32_t wave_index() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t readlane(uint32_t x, int lane) {
    return __builtin_amdgcn_readlane(x, lane);
}

// Inclusive scan over the 64 lanes of a wave with DPP: four row_shr steps
// scan each 16-lane row, then row_bcast:15 and row_bcast:31 carry row totals
// across rows (GFX9 DPP; each step folds into one v_add_u32_dpp).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Sum over the wave, returned uniformly (lane 63 of the inclusive scan).
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) { return readlane(wave_incl_scan(x), 63); }

// ---------------------------------------------------------------------------
// Tile status words.
//
// One 64-bit word per tile (at word kSlotWords * tile: one tile per 128-byte
// slot; packed 8-byte words measured 7% slower on the 2^28 scan, 5690 vs
// 6120 GB/s, as neighbouring tiles' publications from different XCDs
// contended for one cache line), read and written only with relaxed agent-scope
// atomics (global_load/store_dwordx2 ... sc1 on gfx950: coherent across the
// eight XCD L2s without the L2 write-back a release would cost). The payload
// travels inside the word, so no other memory needs ordering:
//   bits  0..31  value: the tile aggregate (FLAG_AGG) or the inclusive prefix
//                through this tile (FLAG_INCL), full 32 bits, wraps mod 2^32
//   bits 32..33  flag: 1 aggregate, 2 inclusive prefix, 3 claimed (the
//                decoupled fallback's "being reduced by a waiter", no value)
//   bits 34..63  epoch of the launch that wrote it (30 bits)
// A word whose epoch differs from the current launch's is "not ready"; the
// array is therefore never cleared between launches.
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1;
constexpr uint64_t kFlagIncl = 2;
constexpr uint64_t kFlagClaim = 3;
constexpr uint32_t kEpochBits = 30;
constexpr uint32_t kEpochMask = (1u << kEpochBits) - 1u;

__device__ __forceinline__ uint64_t make_status(uint32_t value, uint64_t flag, uint32_t epoch) {
    return (uint64_t)value | (flag << 32) | ((uint64_t)epoch << 34);
}
// Flag of a word for this launch's epoch: 0 (not ready: stale, or claimed
// but not yet reduced), kFlagAgg or kFlagIncl.
__device__ __forceinline__ uint32_t status_flag(uint64_t w, uint32_t epoch) {
    const uint32_t f = (uint32_t)((w >> 32) & 3u);
    return ((uint32_t)(w >> 34) == epoch && f != (uint32_t)kFlagClaim) ? f : 0u;
}
__device__ __forceinline__ bool status_claimed(uint64_t w, uint32_t epoch) {
    return (uint32_t)(w >> 34) == epoch && (uint32_t)((w >> 32) & 3u) == (uint32_t)kFlagClaim;
}

__device__ __forceinline__ uint64_t atomic_load_agent(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_agent(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Publication with a returned value: once the wave has waited on the return,
// the word is performed at the device coherence point. Used where later
// output stores must not overtake the publication (in-place fallback safety).
__device__ __forceinline__ uint64_t atomic_xchg_agent(uint64_t* p, uint64_t v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_cas_agent(uint64_t* p, uint64_t expected, uint64_t desired) {
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;  // the value found: == the argument iff the swap happened
}
__device__ __forceinline__ void atomic_add_counter(unsigned long long* p, unsigned long long v) {
    if (p && v) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Ticket word: (epoch << 32) | tickets_taken. Each workgroup takes one ticket
// with a single 64-bit atomic add, which hands it both its tile index (in
// dispatch order, so every predecessor tile is resident or finished: the
// forward-progress argument of decoupled lookback) and the launch epoch. The
// holder of the last ticket knows every other workgroup has already taken
// theirs and rearms the word for the next launch: ticket 0, epoch + 1
// (skipping epochs that are 0 mod 2^30). No memset is needed between launches,
// so the entry points are one kernel launch and are safe under graph capture.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t next_epoch(uint32_t e) {
    uint32_t n = e + 1u;
    if ((n & kEpochMask) == 0u) ++n;
    return n;
}

struct Ticket {
    uint32_t tile;
    uint32_t epoch;  // already masked to kEpochBits
};

// Called by one thread. `total_tickets` is the number of tickets the launch
// takes in all: the tile count for one-tile-per-workgroup kernels, tiles +
// workgroups for persistent kernels (each persistent workgroup ends on one
// ticket past the last tile).
__device__ __forceinline__ Ticket take_ticket(uint64_t* ticket_word, uint32_t total_tickets) {
    const uint64_t old = __hip_atomic_fetch_add(ticket_word, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Ticket t;
    t.tile = (uint32_t)old;
    const uint32_t e = (uint32_t)(old >> 32);
    t.epoch = e & kEpochMask;
    if (t.tile == total_tickets - 1u) atomic_store_agent(ticket_word, (uint64_t)next_epoch(e) << 32);
    return t;
}

// Epoch before `e` (both masked), the inverse of next_epoch: the second
// launch of a split scan runs under the first launch's epoch, after that
// launch's last ticket has re-armed the word.
__device__ __forceinline__ uint32_t prev_epoch(uint32_t e) {
    uint32_t p = (e - 1u) & kEpochMask;
    if (p == 0u) p = kEpochMask;
    return p;
}

// Split schedule (two launches; stall_period P >= 2, tiles >= 2): launch 1
// takes every tile with index % P != 0 in order, launch 2 the deferred tiles
// 0, P, 2P, ... Ticket q of a launch maps to its q-th tile.
__host__ __device__ __forceinline__ uint32_t split_tile_count(uint32_t num_tiles, uint32_t period, uint32_t pass) {
    const uint32_t deferred = (num_tiles + period - 1u) / period;
    return pass == 2u ? deferred : num_tiles - deferred;
}
__device__ __forceinline__ uint32_t split_tile(uint32_t q, uint32_t period, uint32_t pass) {
    return pass == 2u ? q * period : q + q / (period - 1u) + 1u;
}

// Called by one thread of a persistent workgroup that stops early (a watchdog
// fired) while holding ticket `held`: takes tickets until it holds one past
// the last tile, as a normal exit would, so the launch still takes exactly
// `total_tickets` and the holder of the last one re-arms the word. Tiles
// taken here are never scanned; the launch has already raised an error.
__device__ __forceinline__ void drain_tickets(uint64_t* ticket_word, uint64_t held, uint32_t num_tiles,
                                              uint32_t total_tickets) {
    uint64_t t = held;
    while (true) {
        const uint32_t c = (uint32_t)t;
        if (c == total_tickets - 1u) atomic_store_agent(ticket_word, (uint64_t)next_epoch((uint32_t)(t >> 32)) << 32);
        if (c >= num_tiles) return;
        t = __hip_atomic_fetch_add(ticket_word, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Device-side watchdog: a spin that exceeds this many empty polls records an
// error (read back by the host) and gives up instead of spinning forever.
// Every 1024 polls a spinner also reads the error word, so once one
// workgroup has given up the others stop waiting too and the grid drains.
constexpr uint32_t kWatchdogPolls = 1u << 20;

__device__ __forceinline__ void raise_device_error(uint32_t* err, uint32_t code) {
    if (err) __hip_atomic_fetch_or(err, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Call once per empty poll; true when the spinner must give up.
__device__ __forceinline__ bool watchdog_expired(uint32_t& polls, const uint32_t* err) {
    if ((++polls & 1023u) != 0u) return false;
    if (polls > kWatchdogPolls) return true;
    return err && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}
constexpr uint32_t kErrLookbackWatchdog = 1u;
constexpr uint32_t kErrStallWatchdog = 2u;

// ---------------------------------------------------------------------------
// Window lookback (one full wave) for the exclusive prefix of a tile, i.e. the
// sum mod 2^32 of the aggregates of all tiles before it.
// Lane l reads the word of tile (top - l), top starting at tile - 1. The wave
// consumes the contiguous ready run nearest to the tile: if that run contains
// an inclusive word, the prefix is complete; otherwise the run's aggregates
// are added and the window slides below it. A not-ready nearest word is
// polled again. Unlike the reference's warp lookback, which waits until all
// 32 words of its window are ready, the wave never waits on a word it does
// not need. (Wider windows, several words per lane per poll, measured slower
// on MI355X: the polls' latency, not their count, dominates.)
//
// Split into issue() and consume() so a caller can put independent memory
// traffic (the next tile's loads) between a poll and its use.
// ---------------------------------------------------------------------------
#ifndef GPS_POLL_SLEEP
#define GPS_POLL_SLEEP 1  // s_sleep argument between empty polls (x 64 clocks)
#endif

struct Lookback {
    uint32_t prefix;
    int64_t top;
    uint64_t w;  // this lane's word of the last issued window
    uint32_t empty_polls;

    __device__ __forceinline__ void start(uint32_t tile) {
        prefix = 0;
        top = (int64_t)tile - 1;
        empty_polls = 0;
    }
    __device__ __forceinline__ void issue(const uint64_t* status, uint32_t epoch) {
        // Branch-free: every lane loads (lanes below tile 0 reload word 0)
        // and selects, so no divergent region precedes consume()'s ballots.
        const int64_t idx = top - (int64_t)lane_id();
        const uint64_t v = atomic_load_agent(status + kSlotWords * (uint64_t)(idx >= 0 ? idx : 0));
        w = idx >= 0 ? v : make_status(0u, kFlagIncl, epoch);
    }
    // Folds the issued window in; true once `prefix` is complete. Sets
    // `gave_up` (and returns true) when the watchdog fires.
    __device__ __forceinline__ bool consume(uint32_t epoch, uint32_t* err, bool& gave_up) {
        const uint32_t lane = lane_id();
        const uint32_t f = status_flag(w, epoch);
        const uint64_t ready = ballot(f != 0u);
        const uint64_t incl = ballot(f == (uint32_t)kFlagIncl);
        const uint32_t run = (~ready) ? (uint32_t)__builtin_ctzll(~ready) : 64u;
        const uint32_t first_incl = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
        if (first_incl < run) {
            prefix += wave_sum(lane <= first_incl ? (uint32_t)w : 0u);
            return true;
        }
        if (run) {
            prefix += wave_sum(lane < run ? (uint32_t)w : 0u);
            top -= run;
            empty_polls = 0;
        } else {
            if (watchdog_expired(empty_polls, err)) {
                if (lane == 0) raise_device_error(err, kErrLookbackWatchdog);
                gave_up = true;
                return true;
            }
            __builtin_amdgcn_s_sleep(GPS_POLL_SLEEP);
        }
        return false;
    }
    // Polls until complete (after an optional already-issued window).
    __device__ __forceinline__ uint32_t finish(const uint64_t* status, uint32_t epoch, uint32_t* err,
                                               bool issued) {
        bool gave_up = false;
        if (!issued) issue(status, epoch);
        while (!consume(epoch, err, gave_up)) issue(status, epoch);
        return prefix;
    }
};

__device__ __forceinline__ uint32_t lookback_u32(const uint64_t* status, uint32_t tile, uint32_t epoch,
                                                 uint32_t* err) {
    Lookback lb;
    lb.start(tile);
    return lb.finish(status, epoch, err, false);
}

}  // namespace gps
