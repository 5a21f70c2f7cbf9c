// Member-wise scan of 4-member uint32 structs (uint4 elements): for each
// member m independently, out[i].m = sum of in[j].m over j <= i (inclusive)
// or j < i (exclusive), mod 2^32.
//
// Reference counterpart: GPUPrefixSumsWebGPUapis/SharedShaders/csdldf_struct.wgsl:60-300
// (a vec4 struct scanned with one status word per member per tile and a
// per-member lookback). Built here as a chained scan with decoupled lookback:
//   - element (wave w, vector k, lane l) of a tile is struct
//     tile * T + w * 64 * VEC + k * 64 + l, so each lane holds whole structs
//     and every wave instruction moves 1 KiB contiguous;
//   - per vector, four DPP wave scans (one per member) and a run carry per
//     member; wave totals combine through LDS;
//   - a tile publishes four status words (one per member, the integer word
//     layout); a reader takes a tile only when all four carry the same flag
//     and epoch, so a half-updated publication reads as not ready;
//   - persistent (default) or one workgroup per tile, as the plain scan.
#include "gps_internal.hpp"
#include "tile_u32.hpp"

namespace gps {

struct V4Params {
    const uint32_t* in;  // 4 * n uint32, 16-byte aligned
    uint32_t* out;
    uint64_t n;  // structs
    uint32_t num_tiles;
    uint64_t* status;  // 4 words per tile
    uint64_t* ticket;
    uint32_t* err;
    unsigned long long* occupancy;  // persistent kernel: +1 per workgroup that takes a tile, or null
};

__device__ __forceinline__ void publish_v4(uint64_t* status, uint32_t tile, u32x4 v, uint64_t flag, uint32_t epoch) {
#pragma unroll
    for (int m = 0; m < 4; ++m) atomic_store_agent(status + kSlotWords * tile + m, make_status(v[m], flag, epoch));
}

// Window lookback over the four-word tiles: lane l reads member l % 4 of
// tile top - l / 4, so one poll is one 8-byte load per lane and covers 16
// tiles. (Four loads per lane over 64 tiles measured 17% slower at 2^26
// structs: the poll's requests, not its reach, set the pace.) A tile counts
// when its four lanes agree on flag and epoch.
struct LookbackV4 {
    u32x4 prefix;
    int64_t top;
    uint64_t w;
    uint32_t empty_polls;

    __device__ __forceinline__ void start(uint32_t tile) {
        prefix = (u32x4){0u, 0u, 0u, 0u};
        top = (int64_t)tile - 1;
        empty_polls = 0;
    }
    // Branch-free window read (lanes below tile 0 reload tile 0 and take the
    // inclusive sentinel), so no divergent region precedes the ballots.
    __device__ __forceinline__ void issue(const uint64_t* status, uint32_t epoch) {
        const uint32_t lane = lane_id();
        const int64_t idx = top - (int64_t)(lane >> 2);
        const int64_t at = idx >= 0 ? idx : 0;
        const uint64_t v = atomic_load_agent(status + kSlotWords * at + (lane & 3u));
        w = idx >= 0 ? v : make_status(0u, kFlagIncl, epoch);
    }
    // Bit 4t set when all four (any of the four) bits of nibble t are set.
    static __device__ __forceinline__ uint64_t all4(uint64_t x) {
        return x & (x >> 1) & (x >> 2) & (x >> 3) & 0x1111111111111111ull;
    }
    static __device__ __forceinline__ uint64_t any4(uint64_t x) {
        return (x | (x >> 1) | (x >> 2) | (x >> 3)) & 0x1111111111111111ull;
    }
    __device__ __forceinline__ bool consume(uint32_t epoch, uint32_t* err) {
        const uint32_t lane = lane_id();
        const uint32_t f = status_flag(w, epoch);
        const uint64_t ready = ballot(f != 0u);
        const uint64_t incl = ballot(f == (uint32_t)kFlagIncl);
        // A tile is readable when its four words are ready and either all or
        // none of them are inclusive (a publication caught halfway is not).
        const uint64_t i_all = all4(incl);
        const uint64_t ok = all4(ready) & ~(any4(incl) & ~i_all);
        const uint64_t ok_incl = ok & i_all;
        const uint64_t not_ok = ~ok & 0x1111111111111111ull;
        const uint32_t run = not_ok ? (uint32_t)__builtin_ctzll(not_ok) >> 2 : 16u;
        const uint32_t first_incl = ok_incl ? (uint32_t)__builtin_ctzll(ok_incl) >> 2 : 16u;
        const uint32_t take = first_incl < run ? first_incl + 1u : run;
        if (take) {
            const uint32_t v = (lane >> 2) < take ? (uint32_t)w : 0u;
#pragma unroll
            for (int m = 0; m < 4; ++m) prefix[m] += wave_sum((lane & 3u) == (uint32_t)m ? v : 0u);
            if (first_incl < run) return true;
            top -= run;
            empty_polls = 0;
            return false;
        }
        if (watchdog_expired(empty_polls, err)) {
            if (lane == 0) raise_device_error(err, kErrLookbackWatchdog);
            return true;
        }
        __builtin_amdgcn_s_sleep(GPS_POLL_SLEEP);
        return false;
    }
    // Polls until complete (after an optional already-issued window).
    __device__ __forceinline__ u32x4 finish(const uint64_t* status, uint32_t epoch, uint32_t* err, bool issued) {
        if (!issued) issue(status, epoch);
        while (!consume(epoch, err)) issue(status, epoch);
        return prefix;
    }
};

// Tile I/O: element (wave, k, lane) of tile t is struct
// t * BLOCK * VEC + wave * 64 * VEC + k * 64 + lane.
template <int BLOCK, int VEC>
__device__ __forceinline__ uint64_t v4_base(uint32_t tile, uint32_t wave, uint32_t lane) {
    return (uint64_t)tile * ((uint64_t)BLOCK * VEC) + (uint64_t)wave * (64u * VEC) + lane;
}

template <int BLOCK, int VEC>
__device__ __forceinline__ void load_v4(u32x4 (&d)[VEC], const V4Params& p, uint32_t tile, uint32_t wave,
                                        uint32_t lane) {
    const uint64_t base = v4_base<BLOCK, VEC>(tile, wave, lane);
    const u32x4* in4 = reinterpret_cast<const u32x4*>(p.in);
    if ((uint64_t)(tile + 1u) * ((uint64_t)BLOCK * VEC) <= p.n) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) d[k] = __builtin_nontemporal_load(in4 + base + (uint64_t)k * 64u);
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const uint64_t i = base + (uint64_t)k * 64u;
            d[k] = i < p.n ? in4[i] : (u32x4){0u, 0u, 0u, 0u};
        }
    }
}

template <int BLOCK, int VEC>
__device__ __forceinline__ void store_v4(const V4Params& p, const u32x4 (&d)[VEC], u32x4 add, uint32_t tile,
                                         uint32_t wave, uint32_t lane) {
    const uint64_t base = v4_base<BLOCK, VEC>(tile, wave, lane);
    u32x4* out4 = reinterpret_cast<u32x4*>(p.out);
    if ((uint64_t)(tile + 1u) * ((uint64_t)BLOCK * VEC) <= p.n) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) __builtin_nontemporal_store(d[k] + add, out4 + base + (uint64_t)k * 64u);
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const uint64_t i = base + (uint64_t)k * 64u;
            if (i < p.n) out4[i] = d[k] + add;
        }
    }
}

// Scans the tile in registers: per vector and member, a DPP scan across the
// 64 lanes plus the member's carry from the earlier vectors; wave totals meet
// in LDS. Leaves each struct's in-wave prefix in d and returns this wave's
// exclusive prefix inside the tile and the tile total, per member. Contains
// one barrier.
template <int BLOCK, int VEC, bool INCL>
__device__ __forceinline__ void scan_v4(u32x4 (&d)[VEC], uint32_t (*s_wave)[BLOCK / kWave], uint32_t wave,
                                        uint32_t lane, u32x4& wave_prefix, u32x4& tile_total) {
    constexpr uint32_t kWaves = BLOCK / kWave;
    u32x4 carry = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        u32x4 o;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t x = d[k][m];
            const uint32_t s = wave_incl_scan(x);
            o[m] = carry[m] + (INCL ? s : s - x);
            carry[m] += readlane(s, 63);
        }
        d[k] = o;
    }
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m) s_wave[m][wave] = carry[m];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        uint32_t wp, tt;
        combine_waves<kWaves>(s_wave[m], wave, wp, tt);
        wave_prefix[m] = wp;
        tile_total[m] = tt;
    }
}

// One workgroup per tile.
template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) void csdl_u32x4(V4Params p) {
    constexpr uint32_t kWaves = BLOCK / kWave;
    __shared__ uint32_t s_wave[4][kWaves];
    __shared__ uint32_t s_ticket[2];
    __shared__ uint32_t s_prefix[4];
    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    if (threadIdx.x == 0) {
        const Ticket t = take_ticket(p.ticket, p.num_tiles);
        s_ticket[0] = t.tile;
        s_ticket[1] = t.epoch;
    }
    __syncthreads();
    const uint32_t tile = s_ticket[0], epoch = s_ticket[1];
    u32x4 d[VEC];
    load_v4<BLOCK, VEC>(d, p, tile, wave, lane);
    u32x4 wave_prefix, tile_total;
    scan_v4<BLOCK, VEC, INCL>(d, s_wave, wave, lane, wave_prefix, tile_total);

    if (wave == 0) {
        u32x4 prefix = {0u, 0u, 0u, 0u};
        if (tile == 0) {
            if (lane == 0) publish_v4(p.status, 0, tile_total, kFlagIncl, epoch);
        } else {
            if (lane == 0) publish_v4(p.status, tile, tile_total, kFlagAgg, epoch);
            LookbackV4 lb;
            lb.start(tile);
            prefix = lb.finish(p.status, epoch, p.err, false);
            if (lane == 0) publish_v4(p.status, tile, prefix + tile_total, kFlagIncl, epoch);
        }
        if (lane == 0) {
#pragma unroll
            for (int m = 0; m < 4; ++m) s_prefix[m] = prefix[m];
        }
    }
    __syncthreads();
    store_v4<BLOCK, VEC>(p, d, (u32x4){s_prefix[0], s_prefix[1], s_prefix[2], s_prefix[3]} + wave_prefix, tile,
                         wave, lane);
}

// Persistent, software-pipelined form (the schedule of the plain CSDL scan's
// default): resident workgroups own streams of tickets; with tile i scanned
// and its aggregates published, a workgroup takes the next ticket, issues
// that tile's loads, then resolves tile i's lookback and stores tile i. Every
// workgroup ends on one ticket past the last tile, so the ticket word re-arms
// after num_tiles + gridDim.x tickets. Forward progress: a tile is only held
// by a running workgroup, and the lowest tile without published aggregates
// belongs to a workgroup whose current tile already has every predecessor
// published.
template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void csdl_pp_u32x4(V4Params p) {
    constexpr uint32_t kWaves = BLOCK / kWave;
    __shared__ uint32_t s_wave[4][kWaves];
    __shared__ uint32_t s_ticket[3];
    __shared__ uint32_t s_prefix[4];
    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    const uint32_t total_tickets = p.num_tiles + gridDim.x;
    if (threadIdx.x == 0) {
        const Ticket t = take_ticket(p.ticket, total_tickets);
        s_ticket[0] = t.tile;
        s_ticket[1] = t.epoch;
    }
    __syncthreads();
    uint32_t tile = s_ticket[0];
    const uint32_t epoch = s_ticket[1];
    if (tile >= p.num_tiles) return;
    if (threadIdx.x == 0) atomic_add_counter(p.occupancy, 1ull);

    u32x4 a[VEC], b[VEC];
    load_v4<BLOCK, VEC>(a, p, tile, wave, lane);
    while (true) {
        u32x4 wave_prefix, tile_total;
        scan_v4<BLOCK, VEC, INCL>(a, s_wave, wave, lane, wave_prefix, tile_total);
        if (threadIdx.x == 0) {
            publish_v4(p.status, tile, tile_total, tile ? kFlagAgg : kFlagIncl, epoch);
            s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
        }
        __syncthreads();
        const uint32_t next = s_ticket[2];
        const bool has_next = next < p.num_tiles;

        LookbackV4 lb;
        if (wave == 0 && tile != 0) {
            lb.start(tile);
            lb.issue(p.status, epoch);
        }
        if (has_next) load_v4<BLOCK, VEC>(b, p, next, wave, lane);
        if (wave == 0) {
            u32x4 prefix = {0u, 0u, 0u, 0u};
            if (tile != 0) {
                prefix = lb.finish(p.status, epoch, p.err, true);
                if (lane == 0) publish_v4(p.status, tile, prefix + tile_total, kFlagIncl, epoch);
            }
            if (lane == 0) {
#pragma unroll
                for (int m = 0; m < 4; ++m) s_prefix[m] = prefix[m];
            }
        }
        __syncthreads();
        store_v4<BLOCK, VEC>(p, a, (u32x4){s_prefix[0], s_prefix[1], s_prefix[2], s_prefix[3]} + wave_prefix,
                             tile, wave, lane);
        if (!has_next) break;
        for (int k = 0; k < VEC; ++k) a[k] = b[k];  // the prefetched tile becomes the current one
        tile = next;
    }
}

// Tiles of at least 4096 structs, so the four status words per tile fit in
// a handle's max_n / 1024 words for any n <= max_n.
#define GPS_V4_SHAPES(X) \
    X(512, 8)            \
    X(1024, 4)           \
    X(1024, 8)

bool v4_shape_supported(TileShape s) {
#define X(B, V) \
    if (s.block == B && s.vec == V) return true;
    GPS_V4_SHAPES(X)
#undef X
    return false;
}

// Persistent grid: as many workgroups as fit on the device at once (the
// occupancy query is made once per kernel instance and process).
template <int B, int V, bool I>
static uint32_t v4_residency() {
    struct Residency {
        uint32_t workgroups = 0;
        Residency() {
            int dev = 0, cus = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess) return;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, csdl_pp_u32x4<B, V, I>, B, 0) != hipSuccess)
                return;
            workgroups = (uint32_t)(cus > 0 ? cus : 1) * (uint32_t)(per_cu > 0 ? per_cu : 1);
        }
    };
    static const Residency r;
    return r.workgroups;
}

template <int B, int V, bool I>
static hipError_t launch_pp_v4(const V4Params& p, hipStream_t stream, bool grid_all_tiles) {
    const uint32_t resident = v4_residency<B, V, I>();
    if (resident == 0) return hipErrorInvalidDevice;
    const uint32_t grid = (grid_all_tiles || p.num_tiles < resident) ? p.num_tiles : resident;
    hipLaunchKernelGGL((csdl_pp_u32x4<B, V, I>), dim3(grid), dim3(B), 0, stream, p);
    return hipGetLastError();
}

uint32_t v4_resident_workgroups(TileShape s, bool inclusive) {
#define X(B, V) \
    if (s.block == B && s.vec == V) return inclusive ? v4_residency<B, V, true>() : v4_residency<B, V, false>();
    GPS_V4_SHAPES(X)
#undef X
    return 0;
}

hipError_t launch_csdl_u32x4(TileShape s, bool inclusive, bool persistent, const uint32_t* in, uint32_t* out,
                             uint64_t n, uint64_t* status, uint64_t* ticket, uint32_t* err, hipStream_t stream,
                             unsigned long long* occupancy, bool grid_all_tiles) {
    V4Params p{};
    p.in = in;
    p.out = out;
    p.n = n;
    p.num_tiles = (uint32_t)((n + (uint64_t)s.block * s.vec - 1) / ((uint64_t)s.block * s.vec));
    p.status = status;
    p.ticket = ticket;
    p.err = err;
    p.occupancy = occupancy;
#define X(B, V)                                                                                            \
    if (s.block == B && s.vec == V) {                                                                      \
        if (persistent)                                                                                    \
            return inclusive ? launch_pp_v4<B, V, true>(p, stream, grid_all_tiles)                         \
                             : launch_pp_v4<B, V, false>(p, stream, grid_all_tiles);                       \
        if (inclusive)                                                                                     \
            hipLaunchKernelGGL((csdl_u32x4<B, V, true>), dim3(p.num_tiles), dim3(B), 0, stream, p);     \
        else                                                                                               \
            hipLaunchKernelGGL((csdl_u32x4<B, V, false>), dim3(p.num_tiles), dim3(B), 0, stream, p);    \
        return hipGetLastError();                                                                          \
    }
    GPS_V4_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace gps
