// Segment-head discovery shared by the batched segmented scans (uint32 and
// floating point). Segment s covers [offsets[s], offsets[s+1]); a tile
// covering [a, b) marks every offset that falls inside it in an LDS bitmap
// with one bit per tile element.
#pragma once

#include "device_common.hpp"

namespace gps {

// First index s in [0, S) with offsets[s] >= a (S if none), by a 64-ary
// search: each round one wave samples 64 evenly spaced offsets and keeps the
// bracket around the first sample >= a. One full wave; returns the same
// value in every lane. Loads are branch-free (lanes past the bracket reload
// a valid entry), so no divergent region precedes a ballot.
__device__ __forceinline__ uint64_t find_first_segment(const uint64_t* offsets, uint64_t S, uint64_t a) {
    const uint32_t lane = lane_id();
    uint64_t lo = 0, hi = S;  // answer in [lo, hi]
    while (hi - lo > 64) {
        const uint64_t step = (hi - lo + 63) / 64;
        uint64_t pos = lo + (uint64_t)(lane + 1) * step - 1;
        if (pos > hi - 1) pos = hi - 1;
        const uint64_t hit = ballot(offsets[pos] >= a);
        if (!hit) return hi;  // every sample (including hi - 1) is below a
        const uint32_t f = (uint32_t)__builtin_ctzll(hit);
        uint64_t prev = lo + (uint64_t)f * step - 1;  // pos of lane f - 1
        if (prev > hi - 1) prev = hi - 1;
        const uint64_t new_lo = f == 0 ? lo : prev + 1;
        hi = lo + (uint64_t)(f + 1) * step - 1 < hi - 1 ? lo + (uint64_t)(f + 1) * step - 1 : hi - 1;
        lo = new_lo;
    }
    // lanes past hi reload offsets[lo], which exists since lo <= S
    const uint64_t pos = lo + lane;
    const bool ge = offsets[pos < hi ? pos : lo] >= a;
    const uint64_t hit = ballot(pos < hi && ge);
    return hit ? lo + (uint64_t)__builtin_ctzll(hit) : hi;
}

// Sets bit (o - a) of s_heads for every offset o in [a, b), scanning the
// offsets from index s0 (the first one >= a) with `count` threads (tid in
// [0, count)); bit 0 of tile 0 is always set (position 0 starts a segment).
// s_heads must be zeroed and s0 visible to those threads (barriers are the
// caller's).
__device__ __forceinline__ void mark_heads_by(uint32_t* s_heads, const uint64_t* offsets, uint64_t S, uint64_t s0,
                                              uint64_t a, uint64_t b, bool first_tile, uint32_t tid,
                                              uint32_t count) {
    if (first_tile && tid == 0) atomicOr(&s_heads[0], 1u);
    for (uint64_t s = s0 + tid; s < S; s += 4ull * count) {
        uint64_t o[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t su = s + (uint64_t)u * count;
            o[u] = su < S ? offsets[su] : b;
        }
        bool past = false;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (o[u] < b) {
                const uint32_t rel = (uint32_t)(o[u] - a);
                atomicOr(&s_heads[rel >> 5], 1u << (rel & 31u));
            } else {
                past = true;
            }
        }
        if (past) break;
    }
}

// mark_heads_by with the whole workgroup.
template <int BLOCK>
__device__ __forceinline__ void mark_heads(uint32_t* s_heads, const uint64_t* offsets, uint64_t S, uint64_t s0,
                                           uint64_t a, uint64_t b, bool first_tile) {
    mark_heads_by(s_heads, offsets, S, s0, a, b, first_tile, threadIdx.x, BLOCK);
}

// The 4 head bits of vector k of this lane in a wave's run of the register
// tile: elements e .. e+3, e = run_base + k * 256 + lane * 4 (run_base a
// multiple of 256), i.e. bits e % 32 .. +3 of bitmap word e / 32. With
// `lane_words` = s_heads + run_base / 32 + lane / 8 that word is
// lane_words[8 k], an immediate LDS offset, so a kernel keeps one address
// and one shift per lane instead of one word index per vector.
__device__ __forceinline__ uint32_t run_nibble(const uint32_t* lane_words, int k, uint32_t lane) {
    return (lane_words[8 * k] >> ((lane & 7u) * 4u)) & 0xFu;
}

// Inclusive segmented wave scan of (f, v) under
//     (f1, v1) + (f2, v2) = (f1 | f2, f2 ? v2 : v1 + v2),
// on the plain scan's DPP schedule. Lanes a step does not feed receive
// (0, 0), the identity.
__device__ __forceinline__ void wave_seg_scan_u32(uint32_t& f, uint32_t& v) {
#define GPS_SEG_U32_STEP(CTRL, RM, BC)                                                            \
    {                                                                                             \
        const uint32_t yv = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, RM, 0xf, BC);  \
        const uint32_t yf = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)f, CTRL, RM, 0xf, BC);  \
        v = f ? v : yv + v;                                                                       \
        f |= yf;                                                                                  \
    }
    GPS_SEG_U32_STEP(0x111, 0xf, true)   // row_shr:1
    GPS_SEG_U32_STEP(0x112, 0xf, true)   // row_shr:2
    GPS_SEG_U32_STEP(0x114, 0xf, true)   // row_shr:4
    GPS_SEG_U32_STEP(0x118, 0xf, true)   // row_shr:8
    GPS_SEG_U32_STEP(0x142, 0xa, false)  // row_bcast:15
    GPS_SEG_U32_STEP(0x143, 0xc, false)  // row_bcast:31
#undef GPS_SEG_U32_STEP
}

// Segmented combine of the waves' (flag, total) pairs from LDS by one DPP
// pass (lane q holds wave q's pair): this wave's exclusive prefix and the
// tile's pair. A serial loop over the totals would hold 2 * WAVES more
// registers.
template <int WAVES>
__device__ __forceinline__ void combine_waves_seg(const uint32_t* s_flag, const uint32_t* s_val, uint32_t wave,
                                                  uint32_t& pre_f, uint32_t& pre_v, uint32_t& tile_f,
                                                  uint32_t& tile_v) {
    const uint32_t lane = lane_id();
    const uint32_t q = lane < (uint32_t)WAVES ? lane : 0u;
    uint32_t f = lane < (uint32_t)WAVES ? s_flag[q] : 0u;
    uint32_t v = lane < (uint32_t)WAVES ? s_val[q] : 0u;
    wave_seg_scan_u32(f, v);
    const uint32_t ef = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)f, 0x138, 0xf, 0xf, true);  // wave_shr:1
    const uint32_t ev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, true);
    pre_f = readlane(ef, (int)wave);
    pre_v = readlane(ev, (int)wave);
    tile_f = readlane(f, WAVES - 1);
    tile_v = readlane(v, WAVES - 1);
}

}  // namespace gps
