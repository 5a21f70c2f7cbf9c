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
//     and epoch, so a half-updated publication reads as not ready.
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
};

__device__ __forceinline__ void publish_v4(uint64_t* status, uint32_t tile, u32x4 v, uint64_t flag, uint32_t epoch) {
#pragma unroll
    for (int m = 0; m < 4; ++m) atomic_store_agent(status + 4ull * tile + m, make_status(v[m], flag, epoch));
}

// Window lookback over the four-word tiles (one full wave): the per-member
// exclusive prefix of `tile`.
struct LookbackV4 {
    u32x4 prefix;
    int64_t top;
    uint64_t w[4];  // this lane's four words of the last issued window
    uint32_t empty_polls;

    __device__ __forceinline__ void start(uint32_t tile) {
        prefix = (u32x4){0u, 0u, 0u, 0u};
        top = (int64_t)tile - 1;
        empty_polls = 0;
    }
    // Branch-free window read (lanes below tile 0 reload tile 0 and take the
    // inclusive sentinel), so no divergent region precedes the ballots.
    __device__ __forceinline__ void issue(const uint64_t* status, uint32_t epoch) {
        const int64_t idx = top - (int64_t)lane_id();
        const int64_t at = idx >= 0 ? idx : 0;
        uint64_t v[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) v[m] = atomic_load_agent(status + 4 * at + m);
        const uint64_t sentinel = make_status(0u, kFlagIncl, epoch);
#pragma unroll
        for (int m = 0; m < 4; ++m) w[m] = idx >= 0 ? v[m] : sentinel;
    }
    __device__ __forceinline__ bool consume(uint32_t epoch, uint32_t* err) {
        const uint32_t lane = lane_id();
        const uint32_t f0 = status_flag(w[0], epoch);
        const bool agree = f0 == status_flag(w[1], epoch) && f0 == status_flag(w[2], epoch) &&
                           f0 == status_flag(w[3], epoch);
        const uint32_t f = agree ? f0 : 0u;
        const uint64_t ready = ballot(f != 0u);
        const uint64_t incl = ballot(f == (uint32_t)kFlagIncl);
        const uint32_t run = (~ready) ? (uint32_t)__builtin_ctzll(~ready) : 64u;
        const uint32_t first_incl = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
        const uint32_t take = first_incl < run ? first_incl + 1u : run;
        if (take) {
#pragma unroll
            for (int m = 0; m < 4; ++m) prefix[m] += wave_sum(lane < take ? (uint32_t)w[m] : 0u);
            if (first_incl < run) return true;
            top -= run;
            empty_polls = 0;
            return false;
        }
        if (watchdog_expired(empty_polls, err)) {
            if (lane == 0) raise_device_error(err, kErrLookbackWatchdog);
            return true;
        }
        __builtin_amdgcn_s_sleep(1);
        return false;
    }
    __device__ __forceinline__ u32x4 finish(const uint64_t* status, uint32_t epoch, uint32_t* err) {
        issue(status, epoch);
        while (!consume(epoch, err)) issue(status, epoch);
        return prefix;
    }
};

template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) void csdl_u32x4(V4Params p) {
    constexpr uint32_t kWaves = BLOCK / kWave;
    constexpr uint64_t kTile = (uint64_t)BLOCK * VEC;  // structs
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
    const uint64_t base = (uint64_t)tile * kTile + (uint64_t)wave * (64u * VEC) + lane;  // struct index
    const bool full = (uint64_t)(tile + 1u) * kTile <= p.n;
    const u32x4* in4 = reinterpret_cast<const u32x4*>(p.in);
    u32x4 d[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const uint64_t i = base + (uint64_t)k * 64u;
        if (full)
            d[k] = __builtin_nontemporal_load(in4 + i);
        else
            d[k] = i < p.n ? in4[i] : (u32x4){0u, 0u, 0u, 0u};
    }

    // Wave run: per vector and member, a DPP scan across the 64 lanes plus
    // the member's carry from the earlier vectors.
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
    u32x4 wave_prefix, tile_total;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        uint32_t wp, tt;
        combine_waves<kWaves>(s_wave[m], wave, wp, tt);
        wave_prefix[m] = wp;
        tile_total[m] = tt;
    }

    if (wave == 0) {
        u32x4 prefix = {0u, 0u, 0u, 0u};
        if (tile == 0) {
            if (lane == 0) publish_v4(p.status, 0, tile_total, kFlagIncl, epoch);
        } else {
            if (lane == 0) publish_v4(p.status, tile, tile_total, kFlagAgg, epoch);
            LookbackV4 lb;
            lb.start(tile);
            prefix = lb.finish(p.status, epoch, p.err);
            if (lane == 0) publish_v4(p.status, tile, prefix + tile_total, kFlagIncl, epoch);
        }
        if (lane == 0) {
#pragma unroll
            for (int m = 0; m < 4; ++m) s_prefix[m] = prefix[m];
        }
    }
    __syncthreads();
    const u32x4 add = (u32x4){s_prefix[0], s_prefix[1], s_prefix[2], s_prefix[3]} + wave_prefix;
    u32x4* out4 = reinterpret_cast<u32x4*>(p.out);
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const uint64_t i = base + (uint64_t)k * 64u;
        if (full)
            __builtin_nontemporal_store(d[k] + add, out4 + i);
        else if (i < p.n)
            out4[i] = d[k] + add;
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

hipError_t launch_csdl_u32x4(TileShape s, bool inclusive, const uint32_t* in, uint32_t* out, uint64_t n,
                             uint64_t* status, uint64_t* ticket, uint32_t* err, hipStream_t stream) {
    V4Params p{};
    p.in = in;
    p.out = out;
    p.n = n;
    p.num_tiles = (uint32_t)((n + (uint64_t)s.block * s.vec - 1) / ((uint64_t)s.block * s.vec));
    p.status = status;
    p.ticket = ticket;
    p.err = err;
#define X(B, V)                                                                                            \
    if (s.block == B && s.vec == V) {                                                                      \
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
