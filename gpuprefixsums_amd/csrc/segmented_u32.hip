// Batched segmented uint32 scan in one launch over the concatenated segments.
//
// No reference counterpart (a north-star extension). The nearest reference
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

// Segmented combine of (has_head, value) pairs, left then right.
__device__ __forceinline__ void seg_combine(uint32_t& f, uint32_t& v, uint32_t f2, uint32_t v2) {
    v = f2 ? v2 : v + v2;
    f |= f2;
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

    // Segmented scan of the wave's run (carry-in 0). For each element keep
    // the locally scanned value; first_head = in-run index of the run's first
    // head (kWaveElems if none): elements before it take the run carry-in.
    const uint32_t run_base = wave * C::kWaveElems;
    uint32_t first_head = C::kWaveElems;
    uint32_t carry = 0, carry_flag = 0;
    u32x4 x[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        if (!INCL) x[k] = d[k];
        const uint32_t e = run_base + (uint32_t)k * 256u + lane * 4u;  // tile-relative index of component 0
        const uint32_t nib = head_nibble(s_heads, e);
        // lane-local segmented inclusive scan
        u32x4 v = d[k];
        v.y = (nib & 2u) ? v.y : v.x + v.y;
        v.z = (nib & 4u) ? v.z : v.y + v.z;
        v.w = (nib & 8u) ? v.w : v.z + v.w;
        const uint32_t part = v.w;
        const uint64_t heads = ballot(nib != 0u);
        const uint32_t s = wave_incl_scan(part);
        const uint32_t ex = s - part;  // plain exclusive scan of the partials
        // nearest head lane strictly below this lane
        const uint64_t below = heads & ((1ull << lane) - 1ull);
        const int j = below ? 63 - __builtin_clzll(below) : 0;
        const uint32_t ex_j = __builtin_amdgcn_ds_bpermute(j << 2, (int)ex);
        const uint32_t lane_carry = below ? ex - ex_j : carry + ex;
        const uint32_t first_c = nib ? (uint32_t)__builtin_ctz(nib) : 4u;
        v.x += first_c > 0u ? lane_carry : 0u;
        v.y += first_c > 1u ? lane_carry : 0u;
        v.z += first_c > 2u ? lane_carry : 0u;
        v.w += first_c > 3u ? lane_carry : 0u;
        d[k] = v;
        if (heads && first_head == C::kWaveElems) {
            const uint32_t fl = (uint32_t)__builtin_ctzll(heads);
            const uint32_t fnib = __builtin_amdgcn_readlane(nib, fl);
            first_head = (uint32_t)k * 256u + fl * 4u + (uint32_t)__builtin_ctz(fnib);
        }
        // run carry after this vector: lane 63's final value
        carry = readlane(v.w, 63);
        carry_flag |= heads != 0ull;
    }
    if (lane == 0) {
        s_wave_val[wave] = carry;
        s_wave_flag[wave] = carry_flag;
    }
    __syncthreads();

    // Tile level: segmented prefix of the waves before this one, and the
    // tile's own (flag, value).
    uint32_t pre_f = 0, pre_v = 0, tile_f = 0, tile_v = 0;
#pragma unroll
    for (uint32_t q = 0; q < C::kWaves; ++q) {
        const uint32_t fq = s_wave_flag[q], vq = s_wave_val[q];
        if (q < wave) seg_combine(pre_f, pre_v, fq, vq);
        seg_combine(tile_f, tile_v, fq, vq);
    }
    const bool starts_with_head = (s_heads[0] & 1u) != 0u;
    if (wave == 0) {
        if (tile_f) {
            if (lane == 0) atomic_store_agent(p.status + tile, make_status(tile_v, kFlagIncl, epoch));
        } else if (lane == 0) {
            atomic_store_agent(p.status + tile, make_status(tile_v, tile ? kFlagAgg : kFlagIncl, epoch));
        }
        uint32_t prefix = 0;
        if (tile != 0 && !starts_with_head) {
            prefix = lookback_u32(p.status, tile, epoch, p.err);
            if (!tile_f && lane == 0)
                atomic_store_agent(p.status + tile, make_status(prefix + tile_v, kFlagIncl, epoch));
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    const uint32_t run_in = pre_f ? pre_v : s_prefix + pre_v;
    uint32_t out4[4];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        u32x4 v = d[k];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t idx = (uint32_t)k * 256u + lane * 4u + (uint32_t)c;
            out4[c] = v[c] + (idx < first_head ? run_in : 0u);
            if (!INCL) out4[c] -= x[k][c];
        }
        d[k] = (u32x4){out4[0], out4[1], out4[2], out4[3]};
    }
    if (vec_path)
        store_vec<VEC>(p.out + base, d, 0u);
    else
        store_guarded<VEC>(p.out, d, 0u, base, p.n);
}

#define GPS_SEG_SHAPES(X) \
    X(256, 4)             \
    X(512, 4)             \
    X(1024, 2)            \
    X(1024, 4)

bool seg_shape_supported(TileShape s) {
#define X(B, V) \
    if (s.block == B && s.vec == V) return true;
    GPS_SEG_SHAPES(X)
#undef X
    return false;
}

hipError_t launch_segscan_u32(TileShape s, bool inclusive, const uint32_t* in, uint32_t* out,
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
