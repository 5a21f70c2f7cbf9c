// uint32 device-wide scans for gfx950:
//   csdl_u32    chained scan with decoupled lookback (single pass, 8 B/elem)
//   csdldf_u32  the same with decoupled fallback: bounded spins, a
//               workgroup-wide redundant reduction of the blocking tile and
//               CAS insertion; optional emulated deadlock (stalled tiles)
//   rts_*_u32   reduce-then-scan (reduce, single-workgroup spine, downsweep)
//
// Reference counterparts:
//   GPUPrefixSumsCUDA/ChainedScanDecoupledLookback.cuh:79-153 (CSDL)
//   GPUPrefixSumsCUDA/EmulatedDeadlocking.cuh:26-168 (fallback + emulation)
//   GPUPrefixSumsD3D12/Shaders/ChainedScanDecoupledLookbackDecoupledFallback.hlsl:81-242
//     (the inclusive and exclusive fallback kernels)
//   GPUPrefixSumsCUDA/ReduceThenScan.cuh:15-161 (RTS)
#include "gps_internal.hpp"
#include "tile_u32.hpp"

namespace gps {

// ---------------------------------------------------------------------------
// Shared front half: take a ticket, load the tile, scan it in registers and
// combine the wave runs. Leaves d scanned (without the tile prefix) and sets
// wave_prefix / tile_total.
// ---------------------------------------------------------------------------
template <int BLOCK, int VEC, bool INCL>
struct TileFront {
    using C = TileCfg<BLOCK, VEC>;
    u32x4 d[VEC];
    uint32_t tile, epoch, wave_prefix, tile_total;
    uint64_t base;
    bool vec_path;

    __device__ __forceinline__ void run(const ScanParams& p, uint32_t* s_wave, uint32_t* s_ticket) {
        const uint32_t wave = wave_index();
        if (threadIdx.x == 0) {
            if (p.split_pass) {
                const Ticket t = take_ticket(p.ticket, split_tile_count(p.num_tiles, p.stall_period, p.split_pass));
                s_ticket[0] = split_tile(t.tile, p.stall_period, p.split_pass);
                s_ticket[1] = p.split_pass == 2u ? prev_epoch(t.epoch) : t.epoch;
            } else {
                const Ticket t = take_ticket(p.ticket, p.num_tiles);
                s_ticket[0] = t.tile;
                s_ticket[1] = t.epoch;
            }
        }
        __syncthreads();
        tile = s_ticket[0];
        epoch = s_ticket[1];
        base = thread_base<BLOCK, VEC>(tile);
        vec_path = p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n);
        if (vec_path)
            load_vec<VEC>(d, p.in + base);
        else
            load_guarded<VEC>(d, p.in, base, p.n);
        const uint32_t run_total = wave_scan_run<VEC, INCL>(d);
        if (lane_id() == 0) s_wave[wave] = run_total;
        __syncthreads();
        combine_waves<C::kWaves>(s_wave, wave, wave_prefix, tile_total);
    }

    __device__ __forceinline__ void store(const ScanParams& p, uint32_t add) {
        if (vec_path)
            store_vec<VEC>(p.out + base, d, add);
        else
            store_guarded<VEC>(p.out, d, add, base, p.n);
    }
};

// ---------------------------------------------------------------------------
// CSDL. Wave 0 publishes the tile aggregate, looks back, and publishes the
// inclusive prefix; every thread then adds tile prefix + wave prefix.
// ---------------------------------------------------------------------------
// Occupancy target for the single-pass kernels: 8 waves per SIMD (<= 64
// VGPRs), i.e. 2 x 1024, 4 x 512 or 8 x 256 threads per CU. A workgroup holds
// its loaded tile in registers through the lookback, so resident bytes in
// flight, not ALU, set the achieved bandwidth.
#ifndef GPS_WAVES_PER_EU
#define GPS_WAVES_PER_EU 8
#endif

template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(GPS_WAVES_PER_EU))) void csdl_u32(
    ScanParams p) {
    using C = TileCfg<BLOCK, VEC>;
    __shared__ uint32_t s_wave[C::kWaves];
    __shared__ uint32_t s_ticket[2];
    __shared__ uint32_t s_prefix;

    TileFront<BLOCK, VEC, INCL> f;
    f.run(p, s_wave, s_ticket);

    if (wave_index() == 0) {
        const uint32_t lane = lane_id();
        if (f.tile == 0) {
            if (lane == 0) {
                const uint32_t carry = p.carry_in ? *p.carry_in : 0u;
                atomic_store_agent(p.status, make_status(carry + f.tile_total, kFlagIncl, f.epoch));
                s_prefix = carry;
                if (p.carry_out && p.num_tiles == 1) *p.carry_out = carry + f.tile_total;
            }
        } else {
            if (lane == 0) atomic_store_agent(p.status + kSlotWords * f.tile, make_status(f.tile_total, kFlagAgg, f.epoch));
            const uint32_t prefix = lookback_u32(p.status, f.tile, f.epoch, p.err);
            if (lane == 0) {
                atomic_store_agent(p.status + kSlotWords * f.tile, make_status(prefix + f.tile_total, kFlagIncl, f.epoch));
                s_prefix = prefix;
                if (p.carry_out && f.tile == p.num_tiles - 1) *p.carry_out = prefix + f.tile_total;
            }
        }
    }
    __syncthreads();
    f.store(p, s_prefix + f.wave_prefix);
}

// ---------------------------------------------------------------------------
// Persistent, software-pipelined CSDL (the MI355X default schedule).
//
// A chained scan's workgroup must wait for its tile's exclusive prefix before
// it can store, and on MI355X one status round trip to the cross-XCD
// coherence point under full HBM load costs about as much as a 16K-element
// tile's loads. Instead of idling through that wait, each workgroup here
// owns a stream of tickets: with tile i scanned and its aggregate published,
// it takes ticket i+1 and issues tile i+1's loads, then resolves tile i's
// lookback (the first poll is issued ahead of the prefetch so the in-order
// vmcnt does not make it wait for those loads) and stores tile i. The grid is
// sized to what is resident; every workgroup ends on exactly one ticket past
// the last tile, so the ticket word re-arms after num_tiles + gridDim.x
// tickets. Two register tiles per thread (current + prefetched).
//
// Forward progress: a tile is only ever held by a running workgroup, and the
// lowest tile without a published aggregate belongs to a workgroup whose
// current tile is lower still and already has every predecessor published.
// ---------------------------------------------------------------------------
template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(VEC >= 8 ? 4 : GPS_WAVES_PER_EU))) void
csdl_pp_u32(ScanParams p) {
    using C = TileCfg<BLOCK, VEC>;
    __shared__ uint32_t s_wave[C::kWaves];
    __shared__ uint32_t s_ticket[3];
    __shared__ uint32_t s_prefix;
    __shared__ uint32_t s_carry;
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
    const uint64_t base = thread_base<BLOCK, VEC>(tile);
    if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n))
        load_vec<VEC>(a, p.in + base);
    else
        load_guarded<VEC>(a, p.in, base, p.n);

    while (true) {
        const uint32_t run = wave_scan_run<VEC, INCL>(a);
        if (lane == 0) s_wave[wave] = run;
        __syncthreads();
        uint32_t wave_prefix, tile_total;
        combine_waves<C::kWaves>(s_wave, wave, wave_prefix, tile_total);
        if (threadIdx.x == 0) {
            // Publish first, then take the next ticket: ticket order then
            // follows publication order, so lookbacks find contiguous ready
            // runs. (Taking the ticket before the scan, or a whole tile
            // ahead, measured 8% slower at 2^28.)
            uint32_t v = tile_total;
            if (tile == 0) {
                s_carry = p.carry_in ? *p.carry_in : 0u;
                v += s_carry;
            }
            atomic_store_agent(p.status + kSlotWords * tile, make_status(v, tile ? kFlagAgg : kFlagIncl, epoch));
            s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
        }
        __syncthreads();
        const uint32_t next = s_ticket[2];
        const bool has_next = next < p.num_tiles;

        Lookback lb;
        if (wave == 0 && tile != 0) {
            lb.start(tile);
            lb.issue(p.status, epoch);
        }
        if (has_next) {
            const uint64_t next_base = thread_base<BLOCK, VEC>(next);
            if (p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n))
                load_vec<VEC>(b, p.in + next_base);
            else
                load_guarded<VEC>(b, p.in, next_base, p.n);
        }
        if (wave == 0) {
            uint32_t prefix = tile == 0 ? s_carry : 0u;  // tile 0 starts from the carry-in
            if (tile != 0) {
                prefix = lb.finish(p.status, epoch, p.err, true);
                if (lane == 0)
                    atomic_store_agent(p.status + kSlotWords * tile, make_status(prefix + tile_total, kFlagIncl, epoch));
            }
            if (lane == 0) {
                s_prefix = prefix;
                if (p.carry_out && tile == p.num_tiles - 1) *p.carry_out = prefix + tile_total;
            }
        }
        __syncthreads();
        const uint32_t add = s_prefix + wave_prefix;
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n))
            store_vec<VEC>(p.out + base, a, add);
        else
            store_guarded<VEC>(p.out, a, add, base, p.n);
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) a[k] = b[k];
        tile = next;
    }
}

// ---------------------------------------------------------------------------
// Persistent CSDL with a control wave (GPS_SCHEDULE_CONTROL).
//
// The workgroup is DW data waves plus one control wave. Data waves only
// load, scan and store (current tile + prefetched next tile in registers);
// the control wave owns everything that waits on the device coherence point:
// tickets, publication and the lookback. The two sides meet through LDS
// sequence counters (parity-buffered slots, release/acquire at workgroup
// scope), so a data wave never holds a poll or an atomic in its vmcnt queue
// and never waits on a barrier for them. Tile = DW * 256 * VEC elements.
//
// Per iteration i (tile t_i):
//   data:    scan t_i -> totals, arrived[w] = i+1 | wait ticket_seq > i+1 ->
//            issue loads of t_{i+1} | wait prefix_seq > i -> store t_i
//   control: ticket atomic for t_{i+1} | wait all arrived > i -> publish
//            aggregate of t_i | publish t_{i+1} (ticket_seq = i+2) | lookback
//            of t_i -> publish inclusive, prefix (prefix_seq = i+1)
// Every LDS wait is bounded by the device watchdog (error bit, then exit).
// ---------------------------------------------------------------------------

// Spins until *addr >= v (workgroup-scope acquire); false if the watchdog fires.
__device__ __forceinline__ bool lds_wait_ge(const uint32_t* addr, uint32_t v, uint32_t* err) {
    uint32_t polls = 0;
    while (__hip_atomic_load(addr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < v) {
        if (watchdog_expired(polls, err)) {
            raise_device_error(err, kErrLookbackWatchdog);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

__device__ __forceinline__ void lds_release(uint32_t* addr, uint32_t v) {
    __hip_atomic_store(addr, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int DW, int VEC, bool INCL>
__global__ __launch_bounds__(64 * (DW + 1)) __attribute__((amdgpu_waves_per_eu(VEC >= 8 ? 4 : GPS_WAVES_PER_EU))) void
csdl_cw_u32(ScanParams p) {
    constexpr uint32_t kWaveElems = 256u * VEC;
    constexpr uint64_t kTile = (uint64_t)DW * kWaveElems;
    __shared__ uint32_t s_tot[2][DW];
    __shared__ uint32_t s_arrived[DW];
    __shared__ uint32_t s_ticket[2];
    __shared__ uint32_t s_prefix[2];
    __shared__ uint32_t s_ticket_seq, s_prefix_seq;
    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    const uint32_t total_tickets = p.num_tiles + gridDim.x;
    if (threadIdx.x < DW) s_arrived[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_ticket_seq = 0;
        s_prefix_seq = 0;
    }
    __syncthreads();

    if (wave == DW) {
        // ---- control wave ----
        uint32_t tile = 0, epoch = 0;
        if (lane == 0) {
            const Ticket t = take_ticket(p.ticket, total_tickets);
            tile = t.tile;
            epoch = t.epoch;
            s_ticket[0] = tile;
            lds_release(&s_ticket_seq, 1u);
        }
        tile = readlane(tile, 0);
        epoch = readlane(epoch, 0);
        for (uint32_t i = 0; tile < p.num_tiles; ++i) {
            // the next ticket is taken before this tile's waves arrive, so its
            // round trip overlaps their scan
            uint64_t old = 0;
            if (lane == 0) old = __hip_atomic_fetch_add(p.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bool ok = true;
            if (lane < (uint32_t)DW) ok = lds_wait_ge(&s_arrived[lane], i + 1u, p.err);
            if (ballot(!ok)) {
                // Give up, but keep the ticket count exact so the word re-arms.
                if (lane == 0) drain_tickets(p.ticket, old, p.num_tiles, total_tickets);
                return;
            }
            const uint32_t total = wave_sum(lane < (uint32_t)DW ? s_tot[i & 1u][lane] : 0u);
            if (lane == 0) {
                atomic_store_agent(p.status + kSlotWords * tile, make_status(total, tile ? kFlagAgg : kFlagIncl, epoch));
                const uint32_t next = (uint32_t)old;
                if (next == total_tickets - 1u)
                    atomic_store_agent(p.ticket, (uint64_t)next_epoch((uint32_t)(old >> 32)) << 32);
                s_ticket[(i + 1u) & 1u] = next;
                lds_release(&s_ticket_seq, i + 2u);
            }
            uint32_t prefix = 0;
            if (tile != 0) {
                prefix = lookback_u32(p.status, tile, epoch, p.err);
                if (lane == 0) atomic_store_agent(p.status + kSlotWords * tile, make_status(prefix + total, kFlagIncl, epoch));
            }
            if (lane == 0) {
                s_prefix[i & 1u] = prefix;
                lds_release(&s_prefix_seq, i + 1u);
            }
            tile = readlane(lane == 0 ? (uint32_t)old : 0u, 0);
        }
        return;
    }

    // ---- data waves ----
    if (!lds_wait_ge(&s_ticket_seq, 1u, p.err)) return;
    uint32_t tile = s_ticket[0];
    if (tile >= p.num_tiles) return;
    u32x4 a[VEC], b[VEC];
    {
        const uint64_t base = (uint64_t)tile * kTile + wave * kWaveElems + lane * 4u;
        if (p.vec_ok && (uint64_t)(tile + 1u) * kTile <= p.n)
            load_vec<VEC>(a, p.in + base);
        else
            load_guarded<VEC>(a, p.in, base, p.n);
    }
    for (uint32_t i = 0;; ++i) {
        const uint32_t run = wave_scan_run<VEC, INCL>(a);
        if (lane == 0) {
            s_tot[i & 1u][wave] = run;
            lds_release(&s_arrived[wave], i + 1u);
        }
        if (!lds_wait_ge(&s_ticket_seq, i + 2u, p.err)) return;
        const uint32_t next = s_ticket[(i + 1u) & 1u];
        const bool has_next = next < p.num_tiles;
        if (has_next) {
            const uint64_t nb = (uint64_t)next * kTile + wave * kWaveElems + lane * 4u;
            if (p.vec_ok && (uint64_t)(next + 1u) * kTile <= p.n)
                load_vec<VEC>(b, p.in + nb);
            else
                load_guarded<VEC>(b, p.in, nb, p.n);
        }
        if (!lds_wait_ge(&s_prefix_seq, i + 1u, p.err)) return;
        const uint32_t wp = wave_sum(lane < wave ? s_tot[i & 1u][lane] : 0u);
        const uint32_t add = s_prefix[i & 1u] + wp;
        const uint64_t base = (uint64_t)tile * kTile + wave * kWaveElems + lane * 4u;
        if (p.vec_ok && (uint64_t)(tile + 1u) * kTile <= p.n)
            store_vec<VEC>(p.out + base, a, add);
        else
            store_guarded<VEC>(p.out, a, add, base, p.n);
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) a[k] = b[k];
        tile = next;
    }
}

// ---------------------------------------------------------------------------
// CSDLDF: lookback with decoupled fallback.
//
// Wave 0 polls the status window; after spin_limit consecutive polls without
// progress it claims the nearest not-ready tile j (CAS of the stale word to a
// claim word) and hands j to the workgroup, which reduces tile j's input from
// memory. Lane 0 then CAS-inserts that aggregate over the claim (inclusive if
// j == 0); on a lost race (the owner published meanwhile) it uses the word it
// found. Either way the lookback continues below j, or ends at an inclusive
// word. A waiter that finds j already claimed keeps polling, so one reduction
// serves all waiters; after kClaimPatience polls on a claim it reduces j
// anyway (every waiter still falls back within a bounded number of polls).
//
// Emulated deadlock: with stall_period > 0, tiles with tile % period == 0 that
// have a successor hold back their own publication until some successor has
// inserted their aggregate (bounded by a watchdog). The reference stalls
// !(idx & 7) (period 8) and also stalls a last tile, which would never be
// released; here the last tile never stalls.
//
// In-place safety (in == out): a fallback reads tile j's input. Its final CAS
// only wins while j has not published (a publication replaces a claim word
// too), and j publishes with a returning exchange that it waits on before its
// first output store, so a winning reduction never saw j's output.
// ---------------------------------------------------------------------------

// Polls a waiter makes on one stale word before it reduces that tile itself:
// spin_limit for the tile's nearest successor and spin_limit more for each
// tile in between (at most 8 x spin_limit). Every waiter still falls back
// within a bounded number of polls, but the nearest one usually inserts
// first, so the waiters behind it do not all re-read the same tile.
__device__ __forceinline__ uint32_t fallback_patience(uint32_t spin_limit, uint32_t tile, int64_t top) {
    const int64_t gap = (int64_t)tile - 1 - top;
    return spin_limit * (1u + (uint32_t)(gap < 7 ? gap : 7));
}

// Polls a waiter makes on a claimed word before it reduces that tile anyway.
constexpr uint32_t kClaimPatience = 64;

// Called by all of wave 0 when its patience on the stale word w0 of tile
// `top` has run out (w0 is lane 0's, i.e. the nearest not-ready tile's,
// word). A waiter claims the tile before reducing it, so one redundant
// reduction serves every waiter: it CASes w0 to a claim word (flag 3).
//   - claim won: returns true with `expect` = the claim word, which the
//     final insertion CASes over (it fails only if the owner published);
//   - w0 is already a claim: keeps waiting (returns false) for up to
//     kClaimPatience polls, then returns true with `expect` = w0, so a
//     waiter still falls back after a bounded number of polls;
//   - CAS lost to a publication or another claim: returns false (poll on).
// The caller resets its spin count when false is returned.
__device__ __forceinline__ bool claim_for_fallback(uint64_t* status, int64_t top, uint64_t w0, uint32_t epoch,
                                                   uint32_t& claim_waits, uint64_t& expect) {
    if (status_claimed(w0, epoch)) {
        if (++claim_waits < kClaimPatience) return false;
        claim_waits = 0;
        expect = w0;
        return true;
    }
    const uint64_t claim = make_status(0u, kFlagClaim, epoch);
    uint64_t found = w0;
    if (lane_id() == 0) found = atomic_cas_agent(status + kSlotWords * (uint64_t)top, w0, claim);
    found = ((uint64_t)readlane((uint32_t)(found >> 32), 0) << 32) | readlane((uint32_t)found, 0);
    if (found != w0) return false;
    claim_waits = 0;
    expect = claim;
    return true;
}

template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) void csdldf_u32(ScanParams p) {
    using C = TileCfg<BLOCK, VEC>;
    __shared__ uint32_t s_wave[C::kWaves];
    __shared__ uint32_t s_red[C::kWaves];
    __shared__ uint32_t s_ticket[2];
    __shared__ uint32_t s_prefix;
    __shared__ uint32_t s_poll_done, s_cas_done, s_fb_tile;

    TileFront<BLOCK, VEC, INCL> f;
    f.run(p, s_wave, s_ticket);
    const uint32_t tile = f.tile, epoch = f.epoch;
    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();

    if (threadIdx.x == 0) {
        unsigned long long stall_timeouts = 0;
        if (p.stall_period && !p.split_pass && tile % p.stall_period == 0 && tile + 1u < p.num_tiles) {
            uint32_t polls = 0;
            while (status_flag(atomic_load_agent(p.status + kSlotWords * tile), epoch) == 0u) {
                if (watchdog_expired(polls, p.err)) {
                    stall_timeouts = 1;
                    raise_device_error(p.err, kErrStallWatchdog);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        atomic_xchg_agent(p.status + kSlotWords * tile, make_status(f.tile_total, tile ? kFlagAgg : kFlagIncl, epoch));
        if (p.counters) atomic_add_counter(p.counters + 3, stall_timeouts);
        if (tile == 0) s_prefix = 0;
    }

    if (tile != 0) {
        // Lookback state, meaningful in wave 0 only (uniform there).
        uint32_t prefix = 0, spins = 0, claim_waits = 0;
        int64_t top = (int64_t)tile - 1;
        uint64_t stale = 0;
        unsigned long long c_polls = 0, c_fallbacks = 0, c_inserts = 0;

        while (true) {
            if (wave == 0) {
                uint32_t done = 0;
                while (true) {
                    const int64_t idx = top - (int64_t)lane;
                    // branch-free window read (no divergent region before the ballots)
                    const uint64_t v = atomic_load_agent(p.status + kSlotWords * (uint64_t)(idx >= 0 ? idx : 0));
                    const uint64_t w = idx >= 0 ? v : make_status(0u, kFlagIncl, epoch);
                    const uint32_t fl = status_flag(w, epoch);
                    const uint64_t ready = ballot(fl != 0u);
                    const uint64_t incl = ballot(fl == (uint32_t)kFlagIncl);
                    const uint32_t run = (~ready) ? (uint32_t)__builtin_ctzll(~ready) : 64u;
                    const uint32_t first_incl = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
                    ++c_polls;
                    if (first_incl < run) {
                        prefix += wave_sum(lane <= first_incl ? (uint32_t)w : 0u);
                        done = 1;
                        break;
                    }
                    if (run) {
                        prefix += wave_sum(lane < run ? (uint32_t)w : 0u);
                        top -= run;
                        spins = 0;
                    } else if (++spins >= fallback_patience(p.spin_limit, tile, top)) {
                        const uint64_t w0 = ((uint64_t)readlane((uint32_t)(w >> 32), 0) << 32) | readlane((uint32_t)w, 0);
                        if (claim_for_fallback(p.status, top, w0, epoch, claim_waits, stale)) break;
                        spins = 0;
                        __builtin_amdgcn_s_sleep(1);
                    } else {
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                if (lane == 0) {
                    if (done) {
                        atomic_xchg_agent(p.status + kSlotWords * tile, make_status(prefix + f.tile_total, kFlagIncl, epoch));
                        s_prefix = prefix;
                    } else {
                        s_fb_tile = (uint32_t)top;
                    }
                    s_poll_done = done;
                }
            }
            __syncthreads();
            if (s_poll_done) break;

            // Workgroup-wide redundant reduction of tile j (always a full tile:
            // only the last tile is partial and it is nobody's predecessor).
            const uint32_t j = s_fb_tile;
            {
                u32x4 fd[VEC];
                const uint64_t fbase = thread_base<BLOCK, VEC>(j);
                if (p.vec_ok)
                    load_vec<VEC>(fd, p.in + fbase);
                else
                    load_guarded<VEC>(fd, p.in, fbase, p.n);
                const uint32_t ws = wave_sum(thread_sum<VEC>(fd));
                if (lane == 0) s_red[wave] = ws;
            }
            __syncthreads();
            if (wave == 0) {
                uint32_t fsum = 0;
#pragma unroll
                for (uint32_t q = 0; q < C::kWaves; ++q) fsum += s_red[q];
                const uint64_t mine = make_status(fsum, j ? kFlagAgg : kFlagIncl, epoch);
                uint64_t found = mine;
                if (lane == 0) found = atomic_cas_agent(p.status + kSlotWords * j, stale, mine);
                found = ((uint64_t)readlane((uint32_t)(found >> 32), 0) << 32) | readlane((uint32_t)found, 0);
                ++c_fallbacks;
                uint32_t fl;
                if (found == stale) {  // inserted ours
                    ++c_inserts;
                    fl = j ? (uint32_t)kFlagAgg : (uint32_t)kFlagIncl;
                    found = mine;
                } else {
                    fl = status_flag(found, epoch);
                }
                uint32_t done = 0;
                if (fl != 0u) {
                    prefix += (uint32_t)found;
                    if (fl == (uint32_t)kFlagIncl) {
                        done = 1;
                    } else {
                        top = (int64_t)j - 1;
                    }
                } else {
                    // A lost CAS always returns a current-epoch word (the only
                    // in-launch writes); anything else is a protocol fault.
                    if (lane == 0) raise_device_error(p.err, kErrLookbackWatchdog);
                    done = 1;
                }
                spins = 0;
                if (lane == 0) {
                    if (done) {
                        atomic_xchg_agent(p.status + kSlotWords * tile, make_status(prefix + f.tile_total, kFlagIncl, epoch));
                        s_prefix = prefix;
                    }
                    s_cas_done = done;
                }
            }
            __syncthreads();
            if (s_cas_done) break;
        }
        if (threadIdx.x == 0 && p.counters) {
            atomic_add_counter(p.counters + 0, c_polls);
            atomic_add_counter(p.counters + 1, c_fallbacks);
            atomic_add_counter(p.counters + 2, c_inserts);
        }
    }
    __syncthreads();
    f.store(p, s_prefix + f.wave_prefix);
}

// ---------------------------------------------------------------------------
// Persistent, software-pipelined CSDLDF (GPS_SCHEDULE_PERSISTENT).
//
// csdl_pp_u32's ticket stream and prefetch (tile i's lookback overlaps tile
// i+1's loads) with csdldf_u32's lookback: bounded spins, a workgroup-wide
// redundant reduction of the blocking tile and CAS insertion.
//
// Emulated deadlock: a stalled tile's workgroup holds back the tile's
// publication until a successor has inserted its aggregate. At stall period 2
// it also holds back its next ticket, so the successor tile is held by
// another workgroup, which takes its ticket, publishes and looks back without
// waiting on anyone, reaches the stalled tile, gives up after spin_limit
// polls and inserts. From period 3 the stalled workgroup first takes its next
// ticket, loads that tile and (unless it stalls too) scans it and publishes
// its aggregate, then waits, so the tile it holds never holds up a lookback.
// Progress from period 3: take the lowest stalled tile s. A workgroup's next
// tile is above its current one, and every other stalled tile is at least
// s + 3, so tiles s + 1 and s + 2 (neither stalls) can only be held by s's
// own workgroup (as its next tile, one of them) or by workgroups that are not
// stalled, or be unticketed yet. A workgroup that is not stalled and holds
// either looks back through s and inserts it; an unticketed one is taken by
// a workgroup that is not stalled; the one case left is s + 1 as the stalled
// workgroup's next tile and the last tile, where the stall is skipped.
// stall_period >= 2 (host-checked) keeps tile + 1 unstalled.
//
// In-place safety: as in csdldf_u32, every publication is a returning
// exchange whose result is checked (so waited on) before the barrier that
// precedes the workgroup's output stores.
// ---------------------------------------------------------------------------
// Inclusive publication of tile > 0 by its owner: the word it replaces is
// this launch's aggregate (the owner's own or a fallback's insertion); any
// other word is a protocol fault. Checking it also makes the owner wait for
// the exchange before its output stores.
__device__ __forceinline__ void publish_incl_checked(const ScanParams& p, uint32_t tile, uint32_t value,
                                                     uint32_t epoch) {
    const uint64_t old = atomic_xchg_agent(p.status + kSlotWords * tile, make_status(value, kFlagIncl, epoch));
    if (status_flag(old, epoch) != (uint32_t)kFlagAgg) raise_device_error(p.err, kErrLookbackWatchdog);
}

// Emulated deadlock: tiles tile % stall_period == 0 that have a successor
// (the split schedule's launches defer those tiles instead).
__device__ __forceinline__ bool is_stalled_tile(const ScanParams& p, uint32_t tile) {
    return p.stall_period && !p.split_pass && tile % p.stall_period == 0 && tile + 1u < p.num_tiles;
}

// One thread: waits until a successor has inserted this tile's word, bounded
// by the device watchdog.
__device__ __forceinline__ void stall_wait(const ScanParams& p, uint32_t tile, uint32_t epoch,
                                           unsigned long long* s_count) {
    uint32_t polls = 0;
    while (status_flag(atomic_load_agent(p.status + kSlotWords * tile), epoch) == 0u) {
        if (watchdog_expired(polls, p.err)) {
            ++s_count[3];
            raise_device_error(p.err, kErrStallWatchdog);
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// One thread: the owner's aggregate (tile 0: inclusive) publication, a
// returning exchange. Only the owner writes an inclusive word for a tile > 0.
__device__ __forceinline__ void publish_own_word(const ScanParams& p, uint32_t tile, uint32_t tile_total,
                                                 uint32_t epoch) {
    const uint64_t old =
        atomic_xchg_agent(p.status + kSlotWords * tile, make_status(tile_total, tile ? kFlagAgg : kFlagIncl, epoch));
    if (tile != 0 && status_flag(old, epoch) == (uint32_t)kFlagIncl) raise_device_error(p.err, kErrLookbackWatchdog);
}

template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(VEC >= 8 ? 4 : GPS_WAVES_PER_EU))) void
csdldf_pp_u32(ScanParams p) {
    using C = TileCfg<BLOCK, VEC>;
    __shared__ uint32_t s_wave[C::kWaves];
    __shared__ uint32_t s_red[C::kWaves];
    __shared__ uint32_t s_ticket[3];
    __shared__ uint32_t s_prefix;
    // Separate flags for the poll and the CAS step: a fast wave 0 may write
    // the next step's flag while slower waves still read this one.
    __shared__ uint32_t s_poll_done, s_cas_done, s_fb_tile;
    __shared__ unsigned long long s_count[4];
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

    // Per-workgroup fallback statistics (polls, fallbacks, insertions, stall
    // timeouts), kept in LDS by lane 0 of wave 0 and added to the counters
    // once at the end.
    if (threadIdx.x < 4) s_count[threadIdx.x] = 0;

    u32x4 a[VEC], b[VEC];
    {
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n))
            load_vec<VEC>(a, p.in + base);
        else
            load_guarded<VEC>(a, p.in, base, p.n);
    }

    // Set when the current tile's aggregate was published while the previous
    // tile stalled (stall-time prepublication, below).
    bool prepublished = false;
    while (true) {
        const uint32_t run = wave_scan_run<VEC, INCL>(a);
        if (lane == 0) s_wave[wave] = run;
        __syncthreads();
        uint32_t wave_prefix, tile_total;
        combine_waves<C::kWaves>(s_wave, wave, wave_prefix, tile_total);
        const bool stalls = is_stalled_tile(p, tile);
        uint32_t next;
        bool has_next;
        bool prefetched = false;
        if (stalls && p.stall_period >= 3u) {
            // Stall-time prepublication. The stalled workgroup takes its next
            // ticket before it waits, loads that tile and, unless it stalls
            // too, reduces it and publishes its aggregate, so no waiter is
            // held up by the tile it holds. The stall is skipped when the next
            // tile is this one's successor and the last tile: no other
            // workgroup would be left to release it. (At period 2 the
            // stalled workgroups could hold every successor between them, so
            // period 2 keeps the plain order.)
            if (threadIdx.x == 0) s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
            __syncthreads();
            next = s_ticket[2];
            has_next = next < p.num_tiles;
            if (has_next) {
                const uint64_t next_base = thread_base<BLOCK, VEC>(next);
                if (p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n))
                    load_vec<VEC>(b, p.in + next_base);
                else
                    load_guarded<VEC>(b, p.in, next_base, p.n);
                prefetched = true;
            }
            const bool pre_next = has_next && !is_stalled_tile(p, next);
            if (pre_next) {
                // The next tile's total (its scan, next iteration, gives the
                // same sum mod 2^32, and skips this publication).
                const uint32_t ws = wave_sum(thread_sum<VEC>(b));
                if (lane == 0) s_red[wave] = ws;
                __syncthreads();
                if (wave == 0) {
                    const uint32_t total = wave_sum(lane < (uint32_t)C::kWaves ? s_red[lane] : 0u);
                    if (lane == 0) {
                        const uint64_t old =
                            atomic_xchg_agent(p.status + kSlotWords * next, make_status(total, kFlagAgg, epoch));
                        if (status_flag(old, epoch) == (uint32_t)kFlagIncl)
                            raise_device_error(p.err, kErrLookbackWatchdog);
                    }
                }
            }
            if (threadIdx.x == 0) {
                if (!(next == tile + 1u && next + 1u == p.num_tiles)) stall_wait(p, tile, epoch, s_count);
                publish_own_word(p, tile, tile_total, epoch);
            }
            __syncthreads();
            prepublished = pre_next;
        } else {
            if (threadIdx.x == 0) {
                if (stalls) stall_wait(p, tile, epoch, s_count);
                if (!prepublished) publish_own_word(p, tile, tile_total, epoch);
                s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
            }
            __syncthreads();
            next = s_ticket[2];
            has_next = next < p.num_tiles;
            prepublished = false;
        }

        // Lookback state (wave 0; uniform there). The first window is issued
        // ahead of the prefetch so the in-order vmcnt does not make its use
        // wait for tile i+1's loads.
        uint32_t prefix = 0, spins = 0, claim_waits = 0;
        int32_t top = (int32_t)tile - 1;  // 32-bit: keeps 1024 x 4 within 64 VGPRs
        uint64_t w = 0, stale = 0;
        if (wave == 0 && tile != 0) {
            const int32_t idx = top - (int32_t)lane;
            const uint64_t v = atomic_load_agent(p.status + kSlotWords * (uint64_t)(idx >= 0 ? idx : 0));
            w = idx >= 0 ? v : make_status(0u, kFlagIncl, epoch);
        }
        if (has_next && !prefetched) {
            const uint64_t next_base = thread_base<BLOCK, VEC>(next);
            if (p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n))
                load_vec<VEC>(b, p.in + next_base);
            else
                load_guarded<VEC>(b, p.in, next_base, p.n);
        }

        while (true) {
            if (wave == 0) {
                uint32_t done = tile == 0 ? 1u : 0u;
                while (!done) {
                    const uint32_t fl = status_flag(w, epoch);
                    const uint64_t ready = ballot(fl != 0u);
                    const uint64_t incl = ballot(fl == (uint32_t)kFlagIncl);
                    const uint32_t run_len = (~ready) ? (uint32_t)__builtin_ctzll(~ready) : 64u;
                    const uint32_t first_incl = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
                    if (lane == 0) ++s_count[0];
                    if (first_incl < run_len) {
                        prefix += wave_sum(lane <= first_incl ? (uint32_t)w : 0u);
                        done = 1;
                        break;
                    }
                    if (run_len) {
                        prefix += wave_sum(lane < run_len ? (uint32_t)w : 0u);
                        top -= run_len;
                        spins = 0;
                    } else if (++spins >= fallback_patience(p.spin_limit, tile, top)) {
                        const uint64_t w0 = ((uint64_t)readlane((uint32_t)(w >> 32), 0) << 32) | readlane((uint32_t)w, 0);
                        if (claim_for_fallback(p.status, top, w0, epoch, claim_waits, stale)) break;
                        spins = 0;
                        __builtin_amdgcn_s_sleep(1);
                    } else {
                        __builtin_amdgcn_s_sleep(1);
                    }
                    const int32_t idx = top - (int32_t)lane;
                    const uint64_t v = atomic_load_agent(p.status + kSlotWords * (uint64_t)(idx >= 0 ? idx : 0));
                    w = idx >= 0 ? v : make_status(0u, kFlagIncl, epoch);
                }
                if (lane == 0) {
                    if (done) {
                        if (tile != 0) publish_incl_checked(p, tile, prefix + tile_total, epoch);
                        s_prefix = prefix;
                    } else {
                        s_fb_tile = (uint32_t)top;
                    }
                    s_poll_done = done;
                }
            }
            __syncthreads();
            if (s_poll_done) break;

            // Workgroup-wide redundant reduction of tile j (a full tile: only
            // the last tile is partial and it is nobody's predecessor), FB
            // vectors per round so that it fits beside the two live tiles.
            const uint32_t j = s_fb_tile;
            {
                constexpr int FB = VEC < 2 ? VEC : 2;
                const uint64_t fbase = thread_base<BLOCK, VEC>(j);
                uint32_t tsum = 0;
#pragma unroll 1
                for (int c = 0; c < VEC; c += FB) {
                    u32x4 fd[FB];
                    if (p.vec_ok)
                        load_vec<FB>(fd, p.in + fbase + (uint64_t)c * 256u);
                    else
                        load_guarded<FB>(fd, p.in, fbase + (uint64_t)c * 256u, p.n);
                    tsum += thread_sum<FB>(fd);
                }
                const uint32_t ws = wave_sum(tsum);
                if (lane == 0) s_red[wave] = ws;
            }
            __syncthreads();
            if (wave == 0) {
                const uint32_t fsum = wave_sum(lane < (uint32_t)C::kWaves ? s_red[lane] : 0u);
                const uint64_t mine = make_status(fsum, j ? kFlagAgg : kFlagIncl, epoch);
                uint64_t found = mine;
                if (lane == 0) found = atomic_cas_agent(p.status + kSlotWords * j, stale, mine);
                found = ((uint64_t)readlane((uint32_t)(found >> 32), 0) << 32) | readlane((uint32_t)found, 0);
                if (lane == 0) ++s_count[1];
                uint32_t fl;
                if (found == stale) {  // inserted ours
                    if (lane == 0) ++s_count[2];
                    fl = j ? (uint32_t)kFlagAgg : (uint32_t)kFlagIncl;
                    found = mine;
                } else {
                    fl = status_flag(found, epoch);
                }
                uint32_t done = 0;
                if (fl != 0u) {
                    prefix += (uint32_t)found;
                    if (fl == (uint32_t)kFlagIncl)
                        done = 1;
                    else
                        top = (int32_t)j - 1;
                } else {
                    // A lost CAS returns a current-epoch word (the only
                    // in-launch writes); anything else is a protocol fault.
                    if (lane == 0) raise_device_error(p.err, kErrLookbackWatchdog);
                    done = 1;
                }
                spins = 0;
                if (lane == 0) {
                    if (done) {
                        publish_incl_checked(p, tile, prefix + tile_total, epoch);
                        s_prefix = prefix;
                    }
                    s_cas_done = done;
                }
                if (!done) {  // re-issue the window below j
                    const int32_t idx = top - (int32_t)lane;
                    const uint64_t v = atomic_load_agent(p.status + kSlotWords * (uint64_t)(idx >= 0 ? idx : 0));
                    w = idx >= 0 ? v : make_status(0u, kFlagIncl, epoch);
                }
            }
            __syncthreads();
            if (s_cas_done) break;
        }

        const uint32_t add = s_prefix + wave_prefix;
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n))
            store_vec<VEC>(p.out + base, a, add);
        else
            store_guarded<VEC>(p.out, a, add, base, p.n);
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) a[k] = b[k];
        tile = next;
    }
    if (threadIdx.x == 0 && p.counters) {
#pragma unroll
        for (int q = 0; q < 4; ++q) atomic_add_counter(p.counters + q, s_count[q]);
    }
}

// ---------------------------------------------------------------------------
// Reduce-then-scan. Three launches: per-tile sums, an exclusive scan of the
// sums by one workgroup, and a rescan of every tile with its carry-in. Reads
// the input twice (12 B/elem of traffic against 8 B algorithmic).
// The reference's partial-tile reduce predicate is inverted
// (ReduceThenScan.cuh:36-37); here the last tile is reduced over [.., n).
// ---------------------------------------------------------------------------
template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void rts_reduce_u32(const uint32_t* in, uint64_t n, uint32_t vec_ok,
                                                        uint32_t* tile_sums) {
    using C = TileCfg<BLOCK, VEC>;
    __shared__ uint32_t s_red[C::kWaves];
    const uint32_t tile = blockIdx.x;
    const uint64_t base = thread_base<BLOCK, VEC>(tile);
    u32x4 d[VEC];
    if (vec_ok && tile_is_full<BLOCK, VEC>(tile, n))
        load_vec<VEC>(d, in + base);
    else
        load_guarded<VEC>(d, in, base, n);
    const uint32_t ws = wave_sum(thread_sum<VEC>(d));
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (uint32_t q = 0; q < C::kWaves; ++q) t += s_red[q];
        tile_sums[tile] = t;
    }
}

// One workgroup of 1024 threads; exclusive scan of tile_sums in place, 4096
// values per round with a uniform carry between rounds.
constexpr int kSpineBlock = 1024;
__global__ __launch_bounds__(kSpineBlock) void rts_spine_u32(uint32_t* tile_sums, uint32_t num_tiles) {
    using C = TileCfg<kSpineBlock, 1>;
    __shared__ uint32_t s_wave[C::kWaves];
    const uint32_t wave = threadIdx.x / kWave;
    uint32_t carry = 0;
    for (uint32_t round = 0; round * C::kTile < num_tiles; ++round) {
        const uint64_t base = thread_base<kSpineBlock, 1>(round);
        u32x4 d[1];
        load_guarded<1>(d, tile_sums, base, num_tiles);
        const uint32_t run = wave_scan_run<1, false>(d);
        if (lane_id() == 0) s_wave[wave] = run;
        __syncthreads();
        uint32_t wave_prefix, total;
        combine_waves<C::kWaves>(s_wave, wave, wave_prefix, total);
        store_guarded<1>(tile_sums, d, carry + wave_prefix, base, num_tiles);
        carry += total;
        __syncthreads();
    }
}

template <int BLOCK, int VEC, bool INCL>
__global__ __launch_bounds__(BLOCK) void rts_downsweep_u32(const uint32_t* in, uint32_t* out, uint64_t n,
                                                           uint32_t vec_ok, const uint32_t* tile_prefix) {
    using C = TileCfg<BLOCK, VEC>;
    __shared__ uint32_t s_wave[C::kWaves];
    const uint32_t tile = blockIdx.x;
    const uint32_t wave = threadIdx.x / kWave;
    const uint64_t base = thread_base<BLOCK, VEC>(tile);
    const bool vec_path = vec_ok && tile_is_full<BLOCK, VEC>(tile, n);
    u32x4 d[VEC];
    if (vec_path)
        load_vec<VEC>(d, in + base);
    else
        load_guarded<VEC>(d, in, base, n);
    const uint32_t run = wave_scan_run<VEC, INCL>(d);
    if (lane_id() == 0) s_wave[wave] = run;
    __syncthreads();
    uint32_t wave_prefix, total;
    combine_waves<C::kWaves>(s_wave, wave, wave_prefix, total);
    const uint32_t add = tile_prefix[tile] + wave_prefix;
    if (vec_path)
        store_vec<VEC>(out + base, d, add);
    else
        store_guarded<VEC>(out, d, add, base, n);
}

// ---------------------------------------------------------------------------
// Utility kernels.
// ---------------------------------------------------------------------------
__global__ void reset_status_kernel(uint64_t* status, uint64_t count, uint64_t* ticket) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) status[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 1ull << 32;
}

// UtilityKernels.cuh:15-20 (InitOne), generalised to any value.
__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

// Counter-based splitmix64 (same stream definition as the host generators):
// element i = top `bits` bits of mix(seed + (first + i + 1) * golden gamma).
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void fill_splitmix_u32_kernel(uint32_t* p, uint64_t n, uint64_t seed, uint64_t first, uint32_t shift) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = (uint32_t)(splitmix64_at(seed, first + i) >> shift);
}

// Device unit kernel for the wave64 primitives (the reference's warp
// primitives, GPUPrefixSumsCUDA/Utils.cuh:22-71): every run of 64
// consecutive elements is one wave's lanes, a tail run zero-padded. The loop
// bound is per wave, so all 64 lanes are active at the DPP steps.
// op 0: inclusive scan, 1: run total in every lane, 2: exclusive scan.
__global__ void wave_unit_kernel(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t op) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t lane = lane_id();
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1u)); g < n; g += stride) {
        const uint64_t i = g + lane;
        const uint32_t v = in[i < n ? i : g];
        const uint32_t x = i < n ? v : 0u;
        const uint32_t s = wave_incl_scan(x);
        const uint32_t r = op == 0u ? s : op == 1u ? readlane(s, 63) : s - x;
        if (i < n) out[i] = r;
    }
}

// UtilityKernels.cuh:22-38: count scan[i] != i+1 (inclusive) / != i (exclusive).
__global__ void validate_ones_kernel(const uint32_t* scan, uint64_t n, uint32_t incl,
                                     unsigned long long* errors) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long e = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        e += scan[i] != (uint32_t)i + incl;
    atomic_add_counter(errors, e);
}

// Parallel form of the serial validator (Utility.hlsl:99-125): a sequence is
// the inclusive scan of x iff out[0] == x[0] and out[i] - out[i-1] == x[i]
// (mod 2^32) for i > 0; exclusive: out[0] == 0 and out[i] - out[i-1] == x[i-1].
__global__ void validate_diff_kernel(const uint32_t* in, const uint32_t* scan, uint64_t n, uint32_t incl,
                                     unsigned long long* errors) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long e = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t prev = i ? scan[i - 1] : 0u;
        const uint32_t x = incl ? in[i] : (i ? in[i - 1] : 0u);
        e += (uint32_t)(scan[i] - prev) != x;
    }
    atomic_add_counter(errors, e);
}

// Streaming copy with the scan's tile geometry and access pattern (dwordx4,
// nontemporal): the practical read-n + write-n ceiling the scan is held to,
// and the calibration kernel for HBM byte counters.
template <int BLOCK, int VEC>
__global__ __launch_bounds__(BLOCK) void copy_u32_kernel(const uint32_t* in, uint32_t* out, uint64_t n,
                                                         uint32_t vec_ok) {
    const uint32_t tile = blockIdx.x;
    const uint64_t base = thread_base<BLOCK, VEC>(tile);
    u32x4 d[VEC];
    if (vec_ok && tile_is_full<BLOCK, VEC>(tile, n)) {
        load_vec<VEC>(d, in + base);
        store_vec<VEC>(out + base, d, 0u);
    } else {
        load_guarded<VEC>(d, in, base, n);
        store_guarded<VEC>(out, d, 0u, base, n);
    }
}

hipError_t launch_copy_u32(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t stream) {
    if (!n) return hipSuccess;
    constexpr uint32_t kTile = TileCfg<256, 4>::kTile;
    const uint32_t vec_ok = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    hipLaunchKernelGGL((copy_u32_kernel<256, 4>), dim3((uint32_t)((n + kTile - 1) / kTile)), dim3(256), 0, stream,
                       in, out, n, vec_ok);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Launchers: runtime tile shape -> template instance.
// ---------------------------------------------------------------------------
#define GPS_SHAPES(X) \
    X(256, 2)         \
    X(256, 4)         \
    X(256, 8)         \
    X(512, 2)         \
    X(512, 4)         \
    X(512, 8)         \
    X(1024, 2)        \
    X(1024, 4)        \
    X(1024, 8)

bool shape_supported(TileShape s) {
#define X(B, V) \
    if (s.block == B && s.vec == V) return true;
    GPS_SHAPES(X)
#undef X
    return false;
}

hipError_t launch_csdl_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream) {
#define X(B, V)                                                                          \
    if (s.block == B && s.vec == V) {                                                    \
        if (inclusive)                                                                   \
            hipLaunchKernelGGL((csdl_u32<B, V, true>), dim3(p.num_tiles), dim3(B), 0, stream, p);  \
        else                                                                             \
            hipLaunchKernelGGL((csdl_u32<B, V, false>), dim3(p.num_tiles), dim3(B), 0, stream, p); \
        return hipGetLastError();                                                        \
    }
    GPS_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

// Persistent grid: as many workgroups as fit on the device at once, at most
// one per CU. A second resident workgroup per CU measured slower for every
// shape that allows one at 2^28 (1024 x 4: 5068 against 5396 GB/s; 512 x 8:
// 5084 against 5474); 1024 x 8 fits one per CU anyway.
#ifndef GPS_PP_WG_PER_CU
#define GPS_PP_WG_PER_CU 1
#endif
template <int B, int V, bool I>
static uint32_t pp_residency() {
    const uint32_t r = resident_workgroups<csdl_pp_u32<B, V, I>, B>();
    const uint32_t cap = device_cus() * GPS_PP_WG_PER_CU;
    return cap && r > cap ? cap : r;
}

template <int B, int V, bool I>
static hipError_t launch_pp(const ScanParams& p, hipStream_t stream, bool grid_all_tiles) {
    const uint32_t resident = pp_residency<B, V, I>();
    if (resident == 0) return hipErrorInvalidDevice;
    const uint32_t grid = grid_all_tiles ? p.num_tiles : persistent_grid(p.num_tiles, resident);
    hipLaunchKernelGGL((csdl_pp_u32<B, V, I>), dim3(grid), dim3(B), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_csdl_pp_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream,
                              bool grid_all_tiles) {
#define X(B, V)                                                                                 \
    if (s.block == B && s.vec == V) {                                                           \
        return inclusive ? launch_pp<B, V, true>(p, stream, grid_all_tiles)                     \
                         : launch_pp<B, V, false>(p, stream, grid_all_tiles);                   \
    }
    GPS_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

uint32_t pp_resident_workgroups(TileShape s, bool inclusive) {
#define X(B, V) \
    if (s.block == B && s.vec == V) return inclusive ? pp_residency<B, V, true>() : pp_residency<B, V, false>();
    GPS_SHAPES(X)
#undef X
    return 0;
}

template <int DW, int V, bool I>
static hipError_t launch_cw(const ScanParams& p, hipStream_t stream) {
    struct Residency {
        uint32_t workgroups = 0;
        Residency() {
            int dev = 0, cus = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess) return;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, csdl_cw_u32<DW, V, I>, 64 * (DW + 1), 0) !=
                hipSuccess)
                return;
            workgroups = (uint32_t)(cus > 0 ? cus : 1) * (uint32_t)(per_cu > 0 ? per_cu : 1);
        }
    };
    static const Residency r;
    if (r.workgroups == 0) return hipErrorInvalidDevice;
    const uint32_t grid = p.num_tiles < r.workgroups ? p.num_tiles : r.workgroups;
    hipLaunchKernelGGL((csdl_cw_u32<DW, V, I>), dim3(grid), dim3(64 * (DW + 1)), 0, stream, p);
    return hipGetLastError();
}

uint32_t cw_tile_elems(TileShape s) { return (s.block - 64u) * s.vec * 4u; }

hipError_t launch_csdl_cw_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream) {
#define X(B, V)                                                                                               \
    if (s.block == B && s.vec == V) {                                                                         \
        return inclusive ? launch_cw<B / 64 - 1, V, true>(p, stream) : launch_cw<B / 64 - 1, V, false>(p, stream); \
    }
    GPS_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

static hipError_t launch_csdldf_grid(TileShape s, bool inclusive, const ScanParams& p, uint32_t grid,
                                     hipStream_t stream) {
#define X(B, V)                                                                            \
    if (s.block == B && s.vec == V) {                                                      \
        if (inclusive)                                                                     \
            hipLaunchKernelGGL((csdldf_u32<B, V, true>), dim3(grid), dim3(B), 0, stream, p);  \
        else                                                                               \
            hipLaunchKernelGGL((csdldf_u32<B, V, false>), dim3(grid), dim3(B), 0, stream, p); \
        return hipGetLastError();                                                          \
    }
    GPS_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_csdldf_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream) {
    return launch_csdldf_grid(s, inclusive, p, p.num_tiles, stream);
}

hipError_t launch_csdldf_split_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream) {
    ScanParams q = p;
    if (p.num_tiles < 2u || p.stall_period < 2u) {  // nothing to defer: one ordinary launch
        q.split_pass = 0;
        q.stall_period = 0;
        return launch_csdldf_grid(s, inclusive, q, q.num_tiles, stream);
    }
    q.split_pass = 1;
    hipError_t e = launch_csdldf_grid(s, inclusive, q, split_tile_count(p.num_tiles, p.stall_period, 1u), stream);
    if (e != hipSuccess) return e;
    q.split_pass = 2;
    return launch_csdldf_grid(s, inclusive, q, split_tile_count(p.num_tiles, p.stall_period, 2u), stream);
}

template <int B, int V, bool I>
static hipError_t launch_df_pp(const ScanParams& p, hipStream_t stream) {
    const uint32_t resident = resident_workgroups<csdldf_pp_u32<B, V, I>, B>();
    if (resident == 0) return hipErrorInvalidDevice;
    const uint32_t grid = p.num_tiles < resident ? p.num_tiles : resident;
    hipLaunchKernelGGL((csdldf_pp_u32<B, V, I>), dim3(grid), dim3(B), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_csdldf_pp_u32(TileShape s, bool inclusive, const ScanParams& p, hipStream_t stream) {
#define X(B, V)                                                                                       \
    if (s.block == B && s.vec == V) {                                                                 \
        return inclusive ? launch_df_pp<B, V, true>(p, stream) : launch_df_pp<B, V, false>(p, stream); \
    }
    GPS_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_rts_u32(TileShape s, bool inclusive, const uint32_t* in, uint32_t* out, uint64_t n,
                          uint32_t num_tiles, uint32_t vec_ok, uint32_t* tile_sums, hipStream_t stream) {
#define X(B, V)                                                                                           \
    if (s.block == B && s.vec == V) {                                                                     \
        hipLaunchKernelGGL((rts_reduce_u32<B, V>), dim3(num_tiles), dim3(B), 0, stream, in, n, vec_ok,    \
                           tile_sums);                                                                    \
        hipLaunchKernelGGL(rts_spine_u32, dim3(1), dim3(kSpineBlock), 0, stream, tile_sums, num_tiles);  \
        if (inclusive)                                                                                    \
            hipLaunchKernelGGL((rts_downsweep_u32<B, V, true>), dim3(num_tiles), dim3(B), 0, stream, in,  \
                               out, n, vec_ok, (const uint32_t*)tile_sums);                               \
        else                                                                                              \
            hipLaunchKernelGGL((rts_downsweep_u32<B, V, false>), dim3(num_tiles), dim3(B), 0, stream, in, \
                               out, n, vec_ok, (const uint32_t*)tile_sums);                               \
        return hipGetLastError();                                                                         \
    }
    GPS_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

static inline dim3 grid_for(uint64_t n, uint32_t block) {
    uint64_t g = (n + block - 1) / block;
    if (g > 4096) g = 4096;
    if (g == 0) g = 1;
    return dim3((uint32_t)g);
}

hipError_t launch_reset_status(uint64_t* status, uint64_t count, uint64_t* ticket, hipStream_t stream) {
    hipLaunchKernelGGL(reset_status_kernel, grid_for(count, 256), dim3(256), 0, stream, status, count, ticket);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t value, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, grid_for(n, 256), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

hipError_t launch_fill_splitmix_u32(uint32_t* p, uint64_t n, uint64_t seed, uint64_t first, uint32_t bits,
                                    hipStream_t stream) {
    if (!n) return hipSuccess;
    if (bits < 1 || bits > 32) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fill_splitmix_u32_kernel, grid_for(n, 256), dim3(256), 0, stream, p, n, seed, first,
                       64u - bits);
    return hipGetLastError();
}

hipError_t launch_wave_unit_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t op, hipStream_t stream) {
    if (!n) return hipSuccess;
    if (op > 2u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(wave_unit_kernel, grid_for(n, 256), dim3(256), 0, stream, in, out, n, op);
    return hipGetLastError();
}

hipError_t launch_validate_ones_u32(const uint32_t* scan, uint64_t n, int inclusive,
                                    unsigned long long* errors, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(validate_ones_kernel, grid_for(n, 256), dim3(256), 0, stream, scan, n,
                       inclusive ? 1u : 0u, errors);
    return hipGetLastError();
}

hipError_t launch_validate_diff_u32(const uint32_t* in, const uint32_t* scan, uint64_t n, int inclusive,
                                    unsigned long long* errors, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(validate_diff_kernel, grid_for(n, 256), dim3(256), 0, stream, in, scan, n,
                       inclusive ? 1u : 0u, errors);
    return hipGetLastError();
}

}  // namespace gps
