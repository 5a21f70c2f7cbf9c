// Batched segmented uint32 scan in one launch over the concatenated segments.
//
// No reference counterpart (an extension beside the reference's scans). The nearest reference
// ideas are the persistent ticket loop (GPUPrefixSumsWebGPUapis/SharedShaders/
// TestVariants/csdldf_occ.wgsl:82-292) and per-member lookback
// (SharedShaders/csdldf_struct.wgsl:153-291).
//
// Segment s covers [offsets[s], offsets[s+1]); position 0 always starts a
// segment. The scan is the chained scan with decoupled lookback under the
// segmented operator (f1, v1) + (f2, v2) = (f1 | f2, f2 ? v2 : v1 + v2):
//   - each tile finds the segment heads inside it with a 64-ary wave search
//     over `offsets` and marks them in an LDS bitmap;
//   - the register tile is scanned with the carry reset at heads (per vector:
//     lane-local segmented scan, one DPP scan of the lane partials, and one
//     ds_bpermute to fetch the partial at the nearest head lane);
//   - a tile that contains a head knows its carry-out without its
//     predecessors and publishes it as INCLUSIVE at once, so lookbacks stop
//     at segment starts; a tile without heads publishes its AGGREGATE;
//   - only elements before the tile's first head take the looked-back carry.
// All arithmetic is mod 2^32, bit-exact with a serial per-segment loop.
#include "gps_internal.hpp"
#include "segments.hpp"
#include "tile_u32.hpp"

namespace gps {

struct SegParams {
    const uint32_t* in;
    uint32_t* out;
    const uint64_t* offsets;  // num_segments + 1 entries
    uint64_t num_segments;
    uint64_t n;
    uint32_t num_tiles;
    uint32_t vec_ok;
    uint64_t* status;
    uint64_t* ticket;
    uint32_t* err;
};

// Segmented scan of one wave's run (256 * VEC elements starting at tile
// index run_base) in registers, with carry-in 0, under the tile's head
// bitmap. On return d holds each element's segmented inclusive (INCL) or
// exclusive scan within the run; elements before the run's first head still
// lack the carry from earlier runs and tiles. Outputs: first_head = in-run
// index of the first head (256 * VEC if none), and the run's (flag, total).
// Per vector: a lane-local segmented scan of the 4 components, one DPP scan of
// the lane totals, and one ds_bpermute for the exclusive value at the nearest
// head lane below (the lane carry is a difference of two plain prefixes,
// exact mod 2^32). The exclusive form needs no copy of the input: element c
// takes 0 at a head, else the inclusive value of element c - 1.
template <int VEC, bool INCL>
__device__ __forceinline__ void seg_scan_run(u32x4 (&d)[VEC], const uint32_t* s_heads, uint32_t run_base,
                                             uint32_t& first_head, uint32_t& carry, uint32_t& carry_flag) {
    const uint32_t lane = lane_id();
    constexpr uint32_t kRun = 256u * VEC;
    first_head = kRun;
    carry = 0;
    carry_flag = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const uint32_t nib = run_nibble(s_heads + run_base / 32u + lane / 8u, k, lane);
        const u32x4 x = d[k];
        u32x4 v = x;  // lane-local segmented inclusive scan
        v.y = (nib & 2u) ? v.y : v.x + v.y;
        v.z = (nib & 4u) ? v.z : v.y + v.z;
        v.w = (nib & 8u) ? v.w : v.z + v.w;
        const uint32_t part = v.w;
        const uint64_t heads = ballot(nib != 0u);
        const uint32_t s = wave_incl_scan(part);
        const uint32_t ex = s - part;  // plain exclusive scan of the lane totals
        const uint64_t below = heads & ((1ull << lane) - 1ull);  // head lanes strictly below
        const int j = below ? 63 - __builtin_clzll(below) : 0;
        const uint32_t ex_j = __builtin_amdgcn_ds_bpermute(j << 2, (int)ex);
        const uint32_t lane_carry = below ? ex - ex_j : carry + ex;
        const uint32_t first_c = nib ? (uint32_t)__builtin_ctz(nib) : 4u;
        u32x4 o;
        if (INCL) {
            o = v;
        } else {
            o.x = 0u;
            o.y = (nib & 2u) ? 0u : v.x;
            o.z = (nib & 4u) ? 0u : v.y;
            o.w = (nib & 8u) ? 0u : v.z;
        }
        o.x += first_c > 0u ? lane_carry : 0u;
        o.y += first_c > 1u ? lane_carry : 0u;
        o.z += first_c > 2u ? lane_carry : 0u;
        o.w += first_c > 3u ? lane_carry : 0u;
        d[k] = o;
        if (heads && first_head == kRun) {
            const uint32_t fl = (uint32_t)__builtin_ctzll(heads);
            const uint32_t fnib = __builtin_amdgcn_readlane(nib, fl);
            first_head = (uint32_t)k * 256u + fl * 4u + (uint32_t)__builtin_ctz(fnib);
        }
        // run carry after this vector: lane 63's segmented inclusive value
        const uint32_t last = v.w + (first_c > 3u ? lane_carry : 0u);
        carry = readlane(last, 63);
        carry_flag |= heads != 0ull;
    }
}

// Adds the carry from earlier runs and tiles to the elements before the
// run's first head.
template <int VEC>
__device__ __forceinline__ void seg_add_carry(u32x4 (&d)[VEC], uint32_t first_head, uint32_t run_in) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int k = 0; k < VEC; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t idx = (uint32_t)k * 256u + lane * 4u + (uint32_t)c;
            d[k][c] += idx < first_head ? run_in : 0u;
        }
}

template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) void segscan_u32(SegParams p) {
    using C = TileCfg<BLOCK, VEC>;
    constexpr uint32_t kWords = C::kTile / 32;  // head bitmap words
    __shared__ uint32_t s_heads[kWords];
    __shared__ uint32_t s_wave_val[C::kWaves];
    __shared__ uint32_t s_wave_flag[C::kWaves];
    __shared__ uint32_t s_ticket[2];
    __shared__ uint32_t s_prefix;
    __shared__ uint64_t s_first_seg;

    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    if (threadIdx.x == 0) {
        const Ticket t = take_ticket(p.ticket, p.num_tiles);
        s_ticket[0] = t.tile;
        s_ticket[1] = t.epoch;
    }
    for (uint32_t i = threadIdx.x; i < kWords; i += BLOCK) s_heads[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_ticket[0], epoch = s_ticket[1];
    const uint64_t a = (uint64_t)tile * C::kTile;
    const uint64_t b = a + C::kTile < p.n ? a + C::kTile : p.n;

    // Data loads first; the head search overlaps their latency.
    const uint64_t base = thread_base<BLOCK, VEC>(tile);
    u32x4 d[VEC];
    const bool vec_path = p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n);
    if (vec_path)
        load_vec<VEC>(d, p.in + base);
    else
        load_guarded<VEC>(d, p.in, base, p.n);

    if (wave == 0) {
        const uint64_t s0 = find_first_segment(p.offsets, p.num_segments, a);
        if (lane == 0) s_first_seg = s0;
    }
    __syncthreads();
    mark_heads<BLOCK>(s_heads, p.offsets, p.num_segments, s_first_seg, a, b, tile == 0);
    __syncthreads();

    uint32_t first_head, carry, carry_flag;
    seg_scan_run<VEC, INCL>(d, s_heads, wave * C::kWaveElems, first_head, carry, carry_flag);
    if (lane == 0) {
        s_wave_val[wave] = carry;
        s_wave_flag[wave] = carry_flag;
    }
    __syncthreads();

    // Tile level: segmented prefix of the waves before this one, and the
    // tile's own (flag, value).
    uint32_t pre_f, pre_v, tile_f, tile_v;
    combine_waves_seg<C::kWaves>(s_wave_flag, s_wave_val, wave, pre_f, pre_v, tile_f, tile_v);
    const bool starts_with_head = (s_heads[0] & 1u) != 0u;
    if (wave == 0) {
        if (tile_f) {
            if (lane == 0) atomic_store_agent(p.status + kSlotWords * tile, make_status(tile_v, kFlagIncl, epoch));
        } else if (lane == 0) {
            atomic_store_agent(p.status + kSlotWords * tile, make_status(tile_v, tile ? kFlagAgg : kFlagIncl, epoch));
        }
        uint32_t prefix = 0;
        if (tile != 0 && !starts_with_head) {
            prefix = lookback_u32(p.status, tile, epoch, p.err);
            if (!tile_f && lane == 0)
                atomic_store_agent(p.status + kSlotWords * tile, make_status(prefix + tile_v, kFlagIncl, epoch));
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    seg_add_carry<VEC>(d, first_head, pre_f ? pre_v : s_prefix + pre_v);
    if (vec_path)
        store_vec<VEC>(p.out + base, d, 0u);
    else
        store_guarded<VEC>(p.out, d, 0u, base, p.n);
}

// ---------------------------------------------------------------------------
// Persistent, software-pipelined form of the same scan (one launch, grid =
// resident workgroups), on the schedule of the persistent CSDL: scan tile i,
// publish its word, take the next ticket, then overlap tile i's lookback with
// the loads of tile i+1. Segment heads are double-buffered in LDS: while wave
// 0 looks back for tile i, the last wave searches the offsets for tile i+1
// and marks its heads into the other buffer (before its own share of the
// prefetch, so the search's dependent loads do not queue behind it). Every
// workgroup ends on exactly one ticket past the last tile, so the ticket word
// re-arms after num_tiles + gridDim.x tickets.
// ---------------------------------------------------------------------------
template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void segscan_pp_u32(
    SegParams p) {
    using C = TileCfg<BLOCK, VEC>;
    constexpr uint32_t kWords = C::kTile / 32;
    constexpr uint32_t kSearchWave = C::kWaves > 1 ? C::kWaves - 1 : 0;
    __shared__ uint32_t s_heads[2][kWords];
    __shared__ uint32_t s_wave_val[C::kWaves];
    __shared__ uint32_t s_wave_flag[C::kWaves];
    __shared__ uint32_t s_ticket[3];
    __shared__ uint32_t s_prefix;
    __shared__ uint64_t s_first_seg;

    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    const uint32_t total_tickets = p.num_tiles + gridDim.x;
    if (threadIdx.x == 0) {
        const Ticket t = take_ticket(p.ticket, total_tickets);
        s_ticket[0] = t.tile;
        s_ticket[1] = t.epoch;
    }
    for (uint32_t i = threadIdx.x; i < kWords; i += BLOCK) s_heads[0][i] = 0u;
    __syncthreads();
    uint32_t tile = s_ticket[0];
    const uint32_t epoch = s_ticket[1];
    if (tile >= p.num_tiles) return;

    u32x4 a[VEC], b[VEC];
    {
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n))
            load_vec<VEC>(a, p.in + base);
        else
            load_guarded<VEC>(a, p.in, base, p.n);
    }
    {
        const uint64_t lo = (uint64_t)tile * C::kTile;
        const uint64_t hi = lo + C::kTile < p.n ? lo + C::kTile : p.n;
        if (wave == 0) {
            const uint64_t s0 = find_first_segment(p.offsets, p.num_segments, lo);
            if (lane == 0) s_first_seg = s0;
        }
        __syncthreads();
        mark_heads<BLOCK>(s_heads[0], p.offsets, p.num_segments, s_first_seg, lo, hi, tile == 0);
        __syncthreads();
    }

    uint32_t cur = 0;
    while (true) {
        uint32_t first_head, carry, carry_flag;
        seg_scan_run<VEC, INCL>(a, s_heads[cur], wave * C::kWaveElems, first_head, carry, carry_flag);
        if (lane == 0) {
            s_wave_val[wave] = carry;
            s_wave_flag[wave] = carry_flag;
        }
        __syncthreads();
        uint32_t pre_f, pre_v, tile_f, tile_v;
        combine_waves_seg<C::kWaves>(s_wave_flag, s_wave_val, wave, pre_f, pre_v, tile_f, tile_v);
        const bool need_lookback = tile != 0 && (s_heads[cur][0] & 1u) == 0u;
        if (threadIdx.x == 0) {
            // a tile with a head knows its carry-out: publish it inclusive now
            atomic_store_agent(p.status + kSlotWords * tile,
                               make_status(tile_v, (tile_f || tile == 0) ? kFlagIncl : kFlagAgg, epoch));
            s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
        }
        // the other head buffer was last read by the previous tile's scan
        for (uint32_t i = threadIdx.x; i < kWords; i += BLOCK) s_heads[cur ^ 1u][i] = 0u;
        __syncthreads();
        const uint32_t next = s_ticket[2];
        const bool has_next = next < p.num_tiles;

        Lookback lb;
        if (wave == 0 && need_lookback) {
            lb.start(tile);
            lb.issue(p.status, epoch);
        }
        if (has_next && wave == kSearchWave) {
            const uint64_t lo = (uint64_t)next * C::kTile;
            const uint64_t hi = lo + C::kTile < p.n ? lo + C::kTile : p.n;
            const uint64_t s0 = find_first_segment(p.offsets, p.num_segments, lo);
            mark_heads_by(s_heads[cur ^ 1u], p.offsets, p.num_segments, s0, lo, hi, next == 0, lane, kWave);
        }
        if (has_next) {
            const uint64_t next_base = thread_base<BLOCK, VEC>(next);
            if (p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n))
                load_vec<VEC>(b, p.in + next_base);
            else
                load_guarded<VEC>(b, p.in, next_base, p.n);
        }
        if (wave == 0) {
            uint32_t prefix = 0;
            if (need_lookback) {
                prefix = lb.finish(p.status, epoch, p.err, true);
                if (!tile_f && lane == 0)
                    atomic_store_agent(p.status + kSlotWords * tile, make_status(prefix + tile_v, kFlagIncl, epoch));
            }
            if (lane == 0) s_prefix = prefix;
        }
        __syncthreads();  // also completes the next tile's head bitmap
        seg_add_carry<VEC>(a, first_head, pre_f ? pre_v : s_prefix + pre_v);
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n))
            store_vec<VEC>(p.out + base, a, 0u);
        else
            store_guarded<VEC>(p.out, a, 0u, base, p.n);
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) a[k] = b[k];
        tile = next;
        cur ^= 1u;
    }
}

// Shapes of the one-workgroup-per-tile kernel, and of the persistent one,
// which holds two tiles per thread within 128 VGPRs at 4 waves per SIMD. At
// 1024 x 8 the exclusive form spills 2 VGPRs and still measured 11% faster at
// 4096 log-normal segments of 50.9 M elements than 1024 x 4 (0.080 against
// 0.090 ms).
#define GPS_SEG_SHAPES(X) \
    X(256, 4)             \
    X(512, 4)             \
    X(1024, 2)            \
    X(1024, 4)            \
    X(1024, 8)
#define GPS_SEG_PP_SHAPES(X) \
    X(256, 4)                \
    X(512, 4)                \
    X(512, 8)                \
    X(1024, 2)               \
    X(1024, 4)               \
    X(1024, 8)

bool seg_shape_supported(TileShape s, bool persistent) {
#define X(B, V) \
    if (s.block == B && s.vec == V) return true;
    if (persistent) {
        GPS_SEG_PP_SHAPES(X)
    } else {
        GPS_SEG_SHAPES(X)
    }
#undef X
    return false;
}

// Persistent grid: as many workgroups as fit on the device at once (queried
// once per kernel instance and process), at most one per tile.
template <int B, int V, bool I>
static uint32_t seg_pp_residency() {
    struct Residency {
        uint32_t workgroups = 0;
        Residency() {
            int dev = 0, cus = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess) return;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, segscan_pp_u32<B, V, I>, B, 0) != hipSuccess)
                return;
            workgroups = (uint32_t)(cus > 0 ? cus : 1) * (uint32_t)(per_cu > 0 ? per_cu : 1);
        }
    };
    static const Residency r;
    return r.workgroups;
}

template <int B, int V, bool I>
static hipError_t launch_seg_pp(const SegParams& p, hipStream_t stream) {
    const uint32_t resident = seg_pp_residency<B, V, I>();
    if (resident == 0) return hipErrorInvalidDevice;
    const uint32_t grid = persistent_grid(p.num_tiles, resident);
    hipLaunchKernelGGL((segscan_pp_u32<B, V, I>), dim3(grid), dim3(B), 0, stream, p);
    return hipGetLastError();
}

hipError_t prepare_segscan_u32(TileShape s, bool inclusive, bool persistent) {
    if (!persistent) return hipSuccess;
#define X(B, V)                                                                                               \
    if (s.block == B && s.vec == V)                                                                           \
        return (inclusive ? seg_pp_residency<B, V, true>() : seg_pp_residency<B, V, false>()) ? hipSuccess \
                                                                                               : hipErrorInvalidDevice;
    GPS_SEG_PP_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_segscan_u32(TileShape s, bool inclusive, bool persistent, const uint32_t* in, uint32_t* out,
                              const uint64_t* offsets, uint64_t num_segments, uint64_t n, uint32_t vec_ok,
                              uint64_t* status, uint64_t* ticket, uint32_t* err, hipStream_t stream) {
    SegParams p{};
    p.in = in;
    p.out = out;
    p.offsets = offsets;
    p.num_segments = num_segments;
    p.n = n;
    p.num_tiles = (uint32_t)((n + s.tile() - 1) / s.tile());
    p.vec_ok = vec_ok;
    p.status = status;
    p.ticket = ticket;
    p.err = err;
    if (persistent) {
#define X(B, V)                                                                                              \
    if (s.block == B && s.vec == V)                                                                          \
        return inclusive ? launch_seg_pp<B, V, true>(p, stream) : launch_seg_pp<B, V, false>(p, stream);
        GPS_SEG_PP_SHAPES(X)
#undef X
        return hipErrorInvalidValue;
    }
#define X(B, V)                                                                                          \
    if (s.block == B && s.vec == V) {                                                                    \
        if (inclusive)                                                                                   \
            hipLaunchKernelGGL((segscan_u32<B, V, true>), dim3(p.num_tiles), dim3(B), 0, stream, p);  \
        else                                                                                             \
            hipLaunchKernelGGL((segscan_u32<B, V, false>), dim3(p.num_tiles), dim3(B), 0, stream, p); \
        return hipGetLastError();                                                                        \
    }
    GPS_SEG_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace gps
