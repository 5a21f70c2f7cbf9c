// Memory-ceiling probes (include/gps_ceiling.h), built into the benchmark
// library libgps_ceiling.so, never into libgps.so.
//
// The persistent scan moves 4 bytes in and 4 bytes out per element, so a copy
// of the same bytes bounds it from above. Which copy depends on the access
// pattern: these probes move the scan's bytes with the patterns it could take
// (tile shape, one-shot or ticketed persistent, nontemporal, staged through
// LDS) and with the runtime's own copy and fill, so the scan's roofline
// fraction can be read against what this part's HBM delivers to each pattern.
//
// Layout, as in the scan's tiles: a tile is BLOCK * VEC uint4; wave w owns the
// contiguous run of 64 * VEC vectors at w * 64 * VEC, and its vector k is the
// 1 KiB at k * 64 inside that run (one dwordx4 per lane).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

#include <map>
#include <mutex>

#include "../../include/gps_ceiling.h"
#include "handle.hpp"
#include "device_common.hpp"

using gps::fail;
using gps::hip_fail;

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int BLOCK, int VEC>
__device__ __forceinline__ uint64_t vec_index(uint64_t tile, int k) {
    return tile * (uint64_t)(BLOCK * VEC) + (threadIdx.x / 64u) * (64u * VEC) + (uint64_t)k * 64u + gps::lane_id();
}

// One tile per workgroup; grid = tiles.
template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void copy_oneshot(const u32x4* in, u32x4* out) {
    u32x4 d[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) d[k] = __builtin_nontemporal_load(in + vec_index<BLOCK, VEC>(blockIdx.x, k));
#pragma unroll
    for (int k = 0; k < VEC; ++k) __builtin_nontemporal_store(d[k], out + vec_index<BLOCK, VEC>(blockIdx.x, k));
}

// Persistent copy with the scan's schedule: each workgroup takes tile tickets
// from one counter and issues the next tile's loads before storing the
// current one. Every workgroup ends on exactly one ticket >= tiles, so the
// launch takes tiles + gridDim.x tickets; the workgroup holding the last one
// re-arms the counter for the next launch.
template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void copy_ticketed(const u32x4* in, u32x4* out, uint32_t tiles, uint32_t* ticket) {
    __shared__ uint32_t s_ticket;
    const uint32_t last = tiles + gridDim.x - 1u;
    if (threadIdx.x == 0) {
        const uint32_t t = atomicAdd(ticket, 1u);
        if (t == last) atomicExch(ticket, 0u);
        s_ticket = t;
    }
    __syncthreads();
    uint32_t t = s_ticket;
    if (t >= tiles) return;
    u32x4 a[VEC], b[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) a[k] = __builtin_nontemporal_load(in + vec_index<BLOCK, VEC>(t, k));
    while (true) {
        __syncthreads();  // every thread has read s_ticket
        if (threadIdx.x == 0) {
            const uint32_t x = atomicAdd(ticket, 1u);
            if (x == last) atomicExch(ticket, 0u);
            s_ticket = x;
        }
        __syncthreads();
        const uint32_t next = s_ticket;
        if (next < tiles) {
#pragma unroll
            for (int k = 0; k < VEC; ++k) b[k] = __builtin_nontemporal_load(in + vec_index<BLOCK, VEC>(next, k));
        }
#pragma unroll
        for (int k = 0; k < VEC; ++k) __builtin_nontemporal_store(a[k], out + vec_index<BLOCK, VEC>(t, k));
        if (next >= tiles) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) a[k] = b[k];
        t = next;
    }
}

// The ticketed copy plus the status traffic of the persistent scan's tile,
// without its waits: thread 0 publishes the tile's aggregate word before the
// next ticket, wave 0 reads the 64 predecessors' words once (one lookback
// poll; its ballot keeps the loads live) and publishes the inclusive word
// before the stores. Status words are 8 bytes, one per 128-byte slot, written
// and read with relaxed device-scope atomics, as in the scan.
// UNCACHED only tells the kernel traces apart: the words live wherever the
// caller's pointers point (uncached device memory for the _UC kinds).
template <int BLOCK, int VEC, bool UNCACHED>
__global__ __launch_bounds__(BLOCK) void copy_ticketed_status(const u32x4* in, u32x4* out, uint32_t tiles,
                                                              uint32_t* ticket, uint64_t* status) {
    constexpr uint64_t kSlot = 16;  // words per 128-byte slot
    __shared__ uint32_t s_ticket;
    const uint32_t last = tiles + gridDim.x - 1u;
    const uint32_t wave = threadIdx.x / 64u;
    if (threadIdx.x == 0) {
        const uint32_t t = atomicAdd(ticket, 1u);
        if (t == last) atomicExch(ticket, 0u);
        s_ticket = t;
    }
    __syncthreads();
    uint32_t t = s_ticket;
    if (t >= tiles) return;
    u32x4 a[VEC], b[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) a[k] = __builtin_nontemporal_load(in + vec_index<BLOCK, VEC>(t, k));
    while (true) {
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_store(status + kSlot * t, (uint64_t)t | (1ull << 32), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t x = atomicAdd(ticket, 1u);
            if (x == last) atomicExch(ticket, 0u);
            s_ticket = x;
        }
        __syncthreads();
        const uint32_t next = s_ticket;
        uint64_t word = 0;
        if (wave == 0) {
            const int64_t idx = (int64_t)t - 1 - (int64_t)gps::lane_id();
            word = __hip_atomic_load(status + kSlot * (uint64_t)(idx >= 0 ? idx : 0), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        }
        if (next < tiles) {
#pragma unroll
            for (int k = 0; k < VEC; ++k) b[k] = __builtin_nontemporal_load(in + vec_index<BLOCK, VEC>(next, k));
        }
        if (wave == 0) {
            const uint64_t ready = __ballot((word >> 32) != 0u);
            if (gps::lane_id() == 0)
                __hip_atomic_store(status + kSlot * t, (uint64_t)(uint32_t)ready | (2ull << 32), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < VEC; ++k) __builtin_nontemporal_store(a[k], out + vec_index<BLOCK, VEC>(t, k));
        if (next >= tiles) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) a[k] = b[k];
        t = next;
    }
}

// One tile per workgroup, loaded into LDS by direct-to-LDS dwordx4 loads (no
// VGPRs hold the data in flight), then read back and stored. Each wave reads
// back only the slice its own loads wrote, so waiting for its own loads
// (vmcnt 0) is the only synchronisation needed.
template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void copy_via_lds(const u32x4* in, u32x4* out) {
    __shared__ u32x4 lds[BLOCK * VEC];
    const uint32_t w = threadIdx.x / 64u, l = gps::lane_id();
    const uint64_t base = (uint64_t)blockIdx.x * (BLOCK * VEC) + w * (64u * VEC);
#pragma unroll
    for (int k = 0; k < VEC; ++k)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(in + base + k * 64 + l),
                                         (__attribute__((address_space(3))) void*)(lds + w * 64 * VEC + k * 64), 16,
                                         0, 0);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's LDS slice has landed
#pragma unroll
    for (int k = 0; k < VEC; ++k)
        __builtin_nontemporal_store(lds[w * 64 * VEC + k * 64 + l], out + base + k * 64 + l);
}

// Read-only stream: XOR of the tile; the sink word is written only if the XOR
// equals a constant, which keeps the loads live.
template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void read_stream(const u32x4* in, uint32_t* sink) {
    u32x4 acc = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < VEC; ++k) acc ^= __builtin_nontemporal_load(in + vec_index<BLOCK, VEC>(blockIdx.x, k));
    const uint32_t x = acc.x ^ acc.y ^ acc.z ^ acc.w;
    if (x == 0x9e3779b9u) sink[0] = x;
}

// Write-only stream of plain (cached) stores, the fastest write pattern measured.
template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void write_stream(u32x4* out) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const uint64_t i = vec_index<BLOCK, VEC>(blockIdx.x, k);
        out[i] = (u32x4){(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)k, 0x5bd1e995u};
    }
}

struct Kind {
    const char* name;
    uint32_t block, vec;
};

constexpr Kind kKinds[GPS_CEIL_KINDS] = {
    {"copy 64x1", 64, 1},
    {"copy 256x4", 256, 4},
    {"copy 1024x8", 1024, 8},
    {"ticketed copy 1024x4", 1024, 4},
    {"ticketed copy 1024x8", 1024, 8},
    {"copy via LDS 256x4", 256, 4},
    {"read-only 1024x1", 1024, 1},
    {"write-only 256x1", 256, 1},
    {"hipMemcpyAsync", 1, 1},
    {"hipMemsetD32Async", 1, 1},
    {"ticketed copy 1024x4 + status", 1024, 4},
    {"ticketed copy 1024x8 + status", 1024, 8},
    {"ticketed copy 1024x4 + uncached status", 1024, 4},
    {"ticketed copy 1024x8 + uncached status", 1024, 8},
};

uint64_t granule_bytes(int kind) { return (uint64_t)kKinds[kind].block * kKinds[kind].vec * 16u; }

bool uncached_status(int kind) {
    return kind == GPS_CEIL_TICKET_1024X4_STATUS_UC || kind == GPS_CEIL_TICKET_1024X8_STATUS_UC;
}

bool ticketed(int kind) {
    return kind == GPS_CEIL_TICKET_1024X4 || kind == GPS_CEIL_TICKET_1024X8 || kind == GPS_CEIL_TICKET_1024X4_STATUS ||
           kind == GPS_CEIL_TICKET_1024X8_STATUS || uncached_status(kind);
}

bool with_status(int kind) {
    return kind == GPS_CEIL_TICKET_1024X4_STATUS || kind == GPS_CEIL_TICKET_1024X8_STATUS || uncached_status(kind);
}

// Scratch layout: the ticket word alone in the first 256 bytes, then one
// 128-byte status slot per tile for the kinds with status traffic.
constexpr uint64_t kTicketBytes = 256;

uint64_t needed_bytes(int kind, uint64_t bytes) {
    if (!ticketed(kind)) return 0;
    return kTicketBytes + (with_status(kind) ? bytes / granule_bytes(kind) * 128u : 0u);
}

uint64_t scratch_bytes(int kind, uint64_t bytes) { return uncached_status(kind) ? 0 : needed_bytes(kind, bytes); }

// The uncached-status kinds' scratch: one block of uncached device memory per
// device, zeroed when allocated, grown by allocating a larger one. Uncached
// memory is never freed once a kernel has read it (memory allocated after
// such a free has held wrong values on this platform), so a replaced block
// stays allocated.
std::mutex g_uc_mu;
std::map<int, std::pair<void*, uint64_t>> g_uc_blocks;

hipError_t uncached_scratch(uint64_t bytes, void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_uc_mu);
    auto& b = g_uc_blocks[dev];
    if (b.second < bytes) {
        void* p = nullptr;
        e = hipExtMallocWithFlags(&p, bytes, hipDeviceMallocUncached);
        if (e != hipSuccess) return e;
        e = hipMemset(p, 0, bytes);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) return e;
        b = {p, bytes};
    }
    *out = b.first;
    return hipSuccess;
}

}  // namespace

extern "C" const char* gps_ceiling_name(int kind) {
    return kind >= 0 && kind < GPS_CEIL_KINDS ? kKinds[kind].name : "unknown";
}

extern "C" int gps_ceiling_granule(int kind, uint64_t* granule) {
    if (kind < 0 || kind >= GPS_CEIL_KINDS) return fail(GPS_ERR_INVALID, "unknown ceiling probe %d", kind);
    if (!granule) return fail(GPS_ERR_INVALID, "null output");
    *granule = granule_bytes(kind);
    return GPS_OK;
}

extern "C" int gps_ceiling_scratch_bytes(int kind, uint64_t bytes, uint64_t* scratch) {
    if (kind < 0 || kind >= GPS_CEIL_KINDS) return fail(GPS_ERR_INVALID, "unknown ceiling probe %d", kind);
    if (!scratch) return fail(GPS_ERR_INVALID, "null output");
    *scratch = scratch_bytes(kind, bytes);
    return GPS_OK;
}

extern "C" int gps_ceiling_run(int kind, const void* src, void* dst, uint64_t bytes, void* scratch, void* stream) {
    if (kind < 0 || kind >= GPS_CEIL_KINDS) return fail(GPS_ERR_INVALID, "unknown ceiling probe %d", kind);
    const uint64_t g = granule_bytes(kind);
    if (bytes == 0 || bytes % g)
        return fail(GPS_ERR_INVALID, "%s moves whole %llu-byte tiles; %llu bytes given", kKinds[kind].name,
                    (unsigned long long)g, (unsigned long long)bytes);
    const bool reads = kind != GPS_CEIL_WRITE_256X1 && kind != GPS_CEIL_MEMSET;
    if ((reads && !src) || !dst) return fail(GPS_ERR_INVALID, "null buffer");
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u)
        return fail(GPS_ERR_INVALID, "buffers must be 16-byte aligned");
    const uint64_t tiles = bytes / g;
    if (tiles > 0xFFFFFFFFull - 4096u) return fail(GPS_ERR_TOO_LARGE, "too many tiles");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const u32x4* in = static_cast<const u32x4*>(src);
    u32x4* out = static_cast<u32x4*>(dst);
    const dim3 grid((uint32_t)tiles);
    switch (kind) {
        case GPS_CEIL_COPY_64X1:
            hipLaunchKernelGGL((copy_oneshot<64, 1>), grid, dim3(64), 0, s, in, out);
            break;
        case GPS_CEIL_COPY_256X4:
            hipLaunchKernelGGL((copy_oneshot<256, 4>), grid, dim3(256), 0, s, in, out);
            break;
        case GPS_CEIL_COPY_1024X8:
            hipLaunchKernelGGL((copy_oneshot<1024, 8>), grid, dim3(1024), 0, s, in, out);
            break;
        case GPS_CEIL_TICKET_1024X4:
        case GPS_CEIL_TICKET_1024X8:
        case GPS_CEIL_TICKET_1024X4_STATUS:
        case GPS_CEIL_TICKET_1024X8_STATUS:
        case GPS_CEIL_TICKET_1024X4_STATUS_UC:
        case GPS_CEIL_TICKET_1024X8_STATUS_UC: {
            if (uncached_status(kind)) HIP_TRY(uncached_scratch(needed_bytes(kind, bytes), &scratch));
            if (!scratch) return fail(GPS_ERR_INVALID, "the ticketed copies need scratch (gps_ceiling_scratch_bytes)");
            if (reinterpret_cast<uintptr_t>(scratch) & 255u) return fail(GPS_ERR_INVALID, "scratch must be 256-byte aligned");
            const uint32_t cus = gps::device_cus();
            if (cus == 0) return fail(GPS_ERR_HIP, "no device");
            const dim3 pgrid(cus);
            uint32_t* ticket = static_cast<uint32_t*>(scratch);
            uint64_t* status = reinterpret_cast<uint64_t*>(static_cast<char*>(scratch) + kTicketBytes);
            const uint32_t nt = (uint32_t)tiles;
            if (kind == GPS_CEIL_TICKET_1024X4)
                hipLaunchKernelGGL((copy_ticketed<1024, 4>), pgrid, dim3(1024), 0, s, in, out, nt, ticket);
            else if (kind == GPS_CEIL_TICKET_1024X8)
                hipLaunchKernelGGL((copy_ticketed<1024, 8>), pgrid, dim3(1024), 0, s, in, out, nt, ticket);
            else if (kind == GPS_CEIL_TICKET_1024X4_STATUS)
                hipLaunchKernelGGL((copy_ticketed_status<1024, 4, false>), pgrid, dim3(1024), 0, s, in, out, nt, ticket,
                                   status);
            else if (kind == GPS_CEIL_TICKET_1024X8_STATUS)
                hipLaunchKernelGGL((copy_ticketed_status<1024, 8, false>), pgrid, dim3(1024), 0, s, in, out, nt, ticket,
                                   status);
            else if (kind == GPS_CEIL_TICKET_1024X4_STATUS_UC)
                hipLaunchKernelGGL((copy_ticketed_status<1024, 4, true>), pgrid, dim3(1024), 0, s, in, out, nt, ticket,
                                   status);
            else
                hipLaunchKernelGGL((copy_ticketed_status<1024, 8, true>), pgrid, dim3(1024), 0, s, in, out, nt, ticket,
                                   status);
            break;
        }
        case GPS_CEIL_COPY_LDS_256X4:
            hipLaunchKernelGGL((copy_via_lds<256, 4>), grid, dim3(256), 0, s, in, out);
            break;
        case GPS_CEIL_READ_1024X1:
            hipLaunchKernelGGL((read_stream<1024, 1>), grid, dim3(1024), 0, s, in, static_cast<uint32_t*>(dst));
            break;
        case GPS_CEIL_WRITE_256X1:
            hipLaunchKernelGGL((write_stream<256, 1>), grid, dim3(256), 0, s, out);
            break;
        case GPS_CEIL_MEMCPY:
            HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
            return GPS_OK;
        case GPS_CEIL_MEMSET:
            HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(dst), 0x5bd1e995, bytes / 4, s));
            return GPS_OK;
    }
    HIP_TRY(hipGetLastError());
    return GPS_OK;
}
