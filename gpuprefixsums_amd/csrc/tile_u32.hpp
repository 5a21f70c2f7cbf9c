// Register-tile routines for the uint32 scans (per-thread uint4 vectors,
// wave-striped so every load/store instruction moves 1 KiB contiguous).
//
// Reference counterparts: GPUPrefixSumsCUDA/ScanCommon.cuh:12-135
// (Scan{In,Ex}clusive{Full,Partial}, Propagate{Full,Partial}).
//
// Tile geometry for BLOCK threads and VEC uint4 per thread:
//   element of (wave w, vector k, lane l, component c) =
//       tile * TILE + w * (256 * VEC) + k * 256 + l * 4 + c
// so wave w owns a contiguous run of 256*VEC elements, scanned in order k,
// then lane, then component.
#pragma once

#include "device_common.hpp"

namespace gps {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int BLOCK, int VEC>
struct TileCfg {
    static_assert(BLOCK % 64 == 0 && BLOCK <= 1024, "BLOCK must be whole waves");
    static constexpr uint32_t kWaves = BLOCK / 64;
    static constexpr uint32_t kItems = VEC * 4;
    static constexpr uint32_t kTile = BLOCK * kItems;
    static constexpr uint32_t kWaveElems = 64 * kItems;
};

// Element offset of this thread's first vector inside the buffer.
template <int BLOCK, int VEC>
__device__ __forceinline__ uint64_t thread_base(uint32_t tile) {
    using C = TileCfg<BLOCK, VEC>;
    return (uint64_t)tile * C::kTile + (threadIdx.x / kWave) * C::kWaveElems + lane_id() * 4u;
}

// Full tile, 16-byte aligned: one dwordx4 per vector, nontemporal (streamed once).
template <int VEC>
__device__ __forceinline__ void load_vec(u32x4 (&d)[VEC], const uint32_t* p) {
    const u32x4* v = reinterpret_cast<const u32x4*>(p);
#pragma unroll
    for (int k = 0; k < VEC; ++k) d[k] = __builtin_nontemporal_load(v + k * 64);
}

// Partial tile or unaligned buffer: element-wise, zero beyond n.
template <int VEC>
__device__ __forceinline__ void load_guarded(u32x4 (&d)[VEC], const uint32_t* in, uint64_t base, uint64_t n) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
            d[k][c] = i < n ? in[i] : 0u;
        }
    }
}

template <int VEC>
__device__ __forceinline__ void store_vec(uint32_t* p, const u32x4 (&d)[VEC], uint32_t add) {
    u32x4* v = reinterpret_cast<u32x4*>(p);
#pragma unroll
    for (int k = 0; k < VEC; ++k) __builtin_nontemporal_store(d[k] + add, v + k * 64);
}

template <int VEC>
__device__ __forceinline__ void store_guarded(uint32_t* out, const u32x4 (&d)[VEC], uint32_t add, uint64_t base,
                                              uint64_t n) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
            if (i < n) out[i] = d[k][c] + add;
        }
    }
}

// Scans the wave's run in registers. On return d holds, per element, the
// inclusive (INCL) or exclusive scan of the wave's run; returns the run total
// (uniform). Per vector: serial scan of the 4 components, DPP scan of the
// lane totals, and a running carry kept uniform via readlane.
template <int VEC, bool INCL>
__device__ __forceinline__ uint32_t wave_scan_run(u32x4 (&d)[VEC]) {
    uint32_t carry = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        u32x4 v = d[k];
        v.y += v.x;
        v.z += v.y;
        v.w += v.z;
        const uint32_t t = v.w;
        const uint32_t s = wave_incl_scan(t);
        const uint32_t add = carry + (s - t);
        if (INCL) {
            d[k] = v + add;
        } else {
            d[k] = (u32x4){add, v.x + add, v.y + add, v.z + add};
        }
        carry += readlane(s, 63);
    }
    return carry;
}

// Plain sum of the thread's vectors (fallback and reduce passes).
template <int VEC>
__device__ __forceinline__ uint32_t thread_sum(const u32x4 (&d)[VEC]) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k) s += d[k].x + d[k].y + d[k].z + d[k].w;
    return s;
}

// True when the whole tile lies inside [0, n).
template <int BLOCK, int VEC>
__device__ __forceinline__ bool tile_is_full(uint32_t tile, uint64_t n) {
    return (uint64_t)(tile + 1u) * TileCfg<BLOCK, VEC>::kTile <= n;
}

// Exclusive prefix of this wave inside the tile, and the tile total, from the
// per-wave totals in LDS (after a barrier).
// Lane j loads wave j's total and the wave scans them with DPP, so each
// thread needs one LDS read and two registers whatever the block size.
template <int WAVES>
__device__ __forceinline__ void combine_waves(const uint32_t* s_wave, uint32_t wave, uint32_t& wave_prefix,
                                              uint32_t& tile_total) {
    const uint32_t lane = lane_id();
    const uint32_t v = lane < (uint32_t)WAVES ? s_wave[lane] : 0u;
    const uint32_t s = wave_incl_scan(v);
    wave_prefix = __builtin_amdgcn_readlane(s - v, wave);
    tile_total = readlane(s, WAVES - 1);
}

}  // namespace gps
