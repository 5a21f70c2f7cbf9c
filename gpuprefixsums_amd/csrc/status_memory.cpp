// Device memory for a handle's ticket and status words.
//
// The single-pass scans poll status words that workgroups on other XCDs are
// writing. Those words live in uncached device memory (hipDeviceMallocUncached,
// MTYPE UC): on MI355X the relaxed agent-scope polls and stores then cost less
// HBM time than on ordinary device memory. Measured at 2^28 u32 elements, the
// persistent 1024 x 8 scan goes from 6158 to 6286 GB/s and 50.9 M elements
// from 78.0 to 75.2 us; the compensated FP32 scan at 2^27 from 5249 to 5529.
//
// An uncached block is never returned to the runtime once a kernel may have
// read it: on these parts memory allocated after freeing such a block has
// held wrong values. A released block goes to a per-device free list instead,
// and later handles reuse it (smallest free block that fits), so the memory
// held is the peak of concurrently live handles' status sizes, not their sum.
// GPS_STATUS_MEMORY=plain in the environment (read at each handle creation)
// selects ordinary device memory, which is freed normally.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "handle.hpp"

namespace gps {
namespace {

struct Pool {
    std::mutex mu;
    // device -> (bytes -> block), the free uncached blocks
    std::map<int, std::multimap<uint64_t, void*>> free_blocks;
    uint64_t held = 0;       // bytes of uncached blocks ever allocated
    uint64_t free_bytes = 0; // of which on the free lists
};

Pool& pool() {
    static Pool* p = new Pool();  // never destroyed: blocks outlive static destructors
    return *p;
}

bool plain_requested() {
    const char* v = std::getenv("GPS_STATUS_MEMORY");
    return v && std::strcmp(v, "plain") == 0;
}

}  // namespace

hipError_t status_block_acquire(int device, uint64_t bytes, StatusBlock& out) {
    out = StatusBlock{};
    if (plain_requested()) {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return e;
        out = StatusBlock{p, bytes, false};
        return hipSuccess;
    }
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> lock(P.mu);
        auto& fl = P.free_blocks[device];
        auto it = fl.lower_bound(bytes);
        if (it != fl.end()) {
            out = StatusBlock{it->second, it->first, true};
            P.free_bytes -= it->first;
            fl.erase(it);
            return hipSuccess;
        }
    }
    void* p = nullptr;
    if (hipExtMallocWithFlags(&p, bytes, hipDeviceMallocUncached) != hipSuccess) {
        // A device or driver without uncached allocations: ordinary memory,
        // freed normally (the scans are correct in either).
        (void)hipGetLastError();
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return e;
        out = StatusBlock{p, bytes, false};
        return hipSuccess;
    }
    {
        std::lock_guard<std::mutex> lock(P.mu);
        P.held += bytes;
    }
    out = StatusBlock{p, bytes, true};
    return hipSuccess;
}

void status_block_release(int device, const StatusBlock& b) {
    if (!b.ptr) return;
    if (!b.uncached) {
        (void)hipFree(b.ptr);
        return;
    }
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    P.free_blocks[device].emplace(b.bytes, b.ptr);
    P.free_bytes += b.bytes;
}

void status_memory_stats(uint64_t* held, uint64_t* free_bytes) {
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    if (held) *held = P.held;
    if (free_bytes) *free_bytes = P.free_bytes;
}

}  // namespace gps

extern "C" int gps_status_memory(uint64_t* held_bytes, uint64_t* free_bytes) {
    gps::status_memory_stats(held_bytes, free_bytes);
    return GPS_OK;
}
