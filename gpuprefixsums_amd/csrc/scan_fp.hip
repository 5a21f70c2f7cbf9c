// Floating-point inclusive scans (FP32 or BF16 input, FP32 output) by chained
// scan with decoupled lookback, with optional two-sum compensated carries.
//
// No reference counterpart (an extension beside the reference's scans); the chained-scan
// structure is the uint32 one (GPUPrefixSumsCUDA/ChainedScanDecoupledLookback.cuh:79-153).
//
// Compensated mode carries every partial sum above the element level as an
// unevaluated pair (s, c), value s + c, combined with Knuth's TwoSum (the
// exact rounding error of s1 + s2) and Neumaier's accumulation of the error
// terms: lane totals, the DPP wave scan, the run carry across vectors, wave
// totals, tile aggregates, and the inter-tile lookback chain. Each output is
// recomputed from its input with a pair carry and rounded once to FP32, so
// its error is about one FP32 rounding of the exact prefix, independent of n.
// The plain mode is the same kernel with single floats (c == 0).
//
// This file is compiled with -ffp-contract=off: a fused multiply-add would
// break the TwoSum error terms.
#include "gps_internal.hpp"
#include "segments.hpp"
#include "tile_u32.hpp"

namespace gps {

struct ff {
    float s, c;
};

__device__ __forceinline__ ff two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

template <bool COMP>
__device__ __forceinline__ ff ff_add(ff x, ff y) {
    if (!COMP) return {x.s + y.s, 0.f};
    ff t = two_sum(x.s, y.s);
    t.c += x.c + y.c;
    return t;
}

template <bool COMP>
__device__ __forceinline__ ff ff_add1(ff x, float v) {
    if (!COMP) return {x.s + v, 0.f};
    ff t = two_sum(x.s, v);
    t.c += x.c;
    return t;
}

// DPP move of a float (controls and masks must be constants).
#define GPS_DPP_F(x, CTRL, RM, BC) \
    __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, (x)), CTRL, RM, 0xf, BC))

// Inclusive wave scan of pairs (same DPP schedule as the integer scan).
template <bool COMP>
__device__ __forceinline__ ff wave_incl_scan_ff(ff x) {
#define GPS_FF_STEP(CTRL, RM, BC)                                                 \
    {                                                                             \
        const ff y = {GPS_DPP_F(x.s, CTRL, RM, BC), COMP ? GPS_DPP_F(x.c, CTRL, RM, BC) : 0.f}; \
        x = ff_add<COMP>(y, x);                                                   \
    }
    GPS_FF_STEP(0x111, 0xf, true)
    GPS_FF_STEP(0x112, 0xf, true)
    GPS_FF_STEP(0x114, 0xf, true)
    GPS_FF_STEP(0x118, 0xf, true)
    GPS_FF_STEP(0x142, 0xa, false)
    GPS_FF_STEP(0x143, 0xc, false)
#undef GPS_FF_STEP
    return x;
}

__device__ __forceinline__ float readlane_f(float x, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
}

// Status pair for the FP scans: two 64-bit words per tile, both carrying the
// flag and epoch (same layout as the integer words, value = float bits):
// word 0 of tile t's slot = s, word 1 = c. Readers accept a tile only when both words
// are current and agree on the flag; each flag is written once per launch, so
// agreeing words belong to the same publication, and a read that tears between
// two publications shows two flags and counts as not ready.
//
// Both words move in one 16-byte access, one request per tile and lane
// instead of two. The store is a buffer_store_dwordx4 with the cache policy
// of the agent-scope atomic stores (sc1: written through past the XCD L2).
// The load is a volatile global_load_dwordx4, so the compiler re-issues it on
// every poll, and system-scope coherent (sc0 sc1). At 2^27 against two 8-byte
// atomics per access: FP32 pairs 5393-5466 -> 5895-5950 GB/s, BF16 pairs
// 5148-5156 -> 5310-5417, FP32 plain 5754-5803 -> 6090-6136. (A volatile
// store, sc0 sc1, gave 5672-5845 / 5126-5131; an sc1 buffer load instead of
// the volatile one measured the same for FP32 and 3% slower for BF16.)
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_status __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) volatile u64x2 status_pair_t;
constexpr int kCachePolicySc1 = 16;        // buffer instruction cache-policy bits: sc1
constexpr int kBufferRsrcWord3 = 0x00020000;  // raw 32-bit buffer resource, no swizzle

__device__ __forceinline__ void load_pair_words(const uint64_t* slot, uint64_t& w0, uint64_t& w1) {
    const u64x2 v = *(const status_pair_t*)(slot);
    w0 = v.x;
    w1 = v.y;
}

// Called by one lane at a time (the buffer resource is scalar: it takes the
// address of the first active lane).
template <bool COMP>
__device__ __forceinline__ void publish_ff(uint64_t* status, uint32_t tile, ff v, uint64_t flag, uint32_t epoch) {
    const uint64_t w0 = make_status(__builtin_bit_cast(uint32_t, v.s), flag, epoch);
    const uint64_t w1 = make_status(COMP ? __builtin_bit_cast(uint32_t, v.c) : 0u, flag, epoch);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(status + kSlotWords * tile, 0, 16, kBufferRsrcWord3);
    const u32x4_status d = {(uint32_t)w0, (uint32_t)(w0 >> 32), (uint32_t)w1, (uint32_t)(w1 >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(d, rs, 0, 0, kCachePolicySc1);
}

// Window lookback over pair words (one full wave); returns the exclusive
// prefix of `tile` as a pair.
template <bool COMP>
__device__ __forceinline__ ff lookback_ff(const uint64_t* status, uint32_t tile, uint32_t epoch, uint32_t* err) {
    const uint32_t lane = lane_id();
    ff prefix = {0.f, 0.f};
    int64_t top = (int64_t)tile - 1;
    uint32_t empty_polls = 0;
    while (true) {
        // Branch-free window read: lanes below tile 0 reload pair 0 and take
        // the sentinel, so no divergent region precedes the ballots.
        const int64_t idx = top - (int64_t)lane;
        const int64_t at = idx >= 0 ? idx : 0;
        uint64_t ls, lc;
        load_pair_words(status + kSlotWords * at, ls, lc);
        const uint64_t sentinel = make_status(0u, kFlagIncl, epoch);
        const uint64_t ws = idx >= 0 ? ls : sentinel, wc = idx >= 0 ? lc : sentinel;
        const uint32_t fs = status_flag(ws, epoch), fc = status_flag(wc, epoch);
        const uint32_t f = fs == fc ? fs : 0u;
        const uint64_t ready = ballot(f != 0u);
        const uint64_t incl = ballot(f == (uint32_t)kFlagIncl);
        const uint32_t run = (~ready) ? (uint32_t)__builtin_ctzll(~ready) : 64u;
        const uint32_t first_incl = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
        const uint32_t take = first_incl < run ? first_incl + 1u : run;
        if (take) {
            ff v = {0.f, 0.f};
            if (lane < take) v = {__builtin_bit_cast(float, (uint32_t)ws), __builtin_bit_cast(float, (uint32_t)wc)};
            const ff s = wave_incl_scan_ff<COMP>(v);
            const ff tot = {readlane_f(s.s, 63), COMP ? readlane_f(s.c, 63) : 0.f};
            prefix = ff_add<COMP>(tot, prefix);
            if (first_incl < run) return prefix;
            top -= run;
            empty_polls = 0;
        } else {
            if (watchdog_expired(empty_polls, err)) {
                if (lane == 0) raise_device_error(err, kErrLookbackWatchdog);
                return prefix;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

struct FpParams {
    const void* in;  // float or bf16 bits
    float* out;
    uint64_t n;
    uint32_t num_tiles;
    uint32_t vec_ok;
    uint64_t* status;  // 2 words per tile
    uint64_t* ticket;
    uint32_t* err;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));

template <bool BF16, int VEC>
__device__ __forceinline__ void load_fp(f32x4 (&x)[VEC], const void* in, uint64_t base, uint64_t n, bool vec_path) {
    if (vec_path) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            if (BF16) {
                const u16x4 b = __builtin_nontemporal_load(reinterpret_cast<const u16x4*>(
                    static_cast<const uint16_t*>(in) + base + (uint64_t)k * 256u));
#pragma unroll
                for (int c = 0; c < 4; ++c) x[k][c] = __builtin_bit_cast(float, (uint32_t)b[c] << 16);
            } else {
                x[k] = __builtin_nontemporal_load(
                    reinterpret_cast<const f32x4*>(static_cast<const float*>(in) + base + (uint64_t)k * 256u));
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
                float v = 0.f;
                if (i < n) {
                    if (BF16)
                        v = __builtin_bit_cast(float, (uint32_t)static_cast<const uint16_t*>(in)[i] << 16);
                    else
                        v = static_cast<const float*>(in)[i];
                }
                x[k][c] = v;
            }
        }
    }
}

template <int BLOCK, int VEC, bool BF16, bool COMP>
__global__ __launch_bounds__(BLOCK) void scan_fp(FpParams p) {
    using C = TileCfg<BLOCK, VEC>;
    __shared__ float s_ws[C::kWaves], s_wc[C::kWaves];
    __shared__ uint32_t s_ticket[2];
    __shared__ float s_ps, s_pc;
    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    if (threadIdx.x == 0) {
        const Ticket t = take_ticket(p.ticket, p.num_tiles);
        s_ticket[0] = t.tile;
        s_ticket[1] = t.epoch;
    }
    __syncthreads();
    const uint32_t tile = s_ticket[0], epoch = s_ticket[1];
    const uint64_t base = thread_base<BLOCK, VEC>(tile);
    const bool vec_path = p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n);
    f32x4 x[VEC];
    load_fp<BF16, VEC>(x, p.in, base, p.n, vec_path);

    // Per vector: lane total (pair), wave exclusive scan -> lane carry L_k.
    ff lane_carry[VEC];
    ff run = {0.f, 0.f};
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        ff t = {x[k].x, 0.f};
        t = ff_add1<COMP>(t, x[k].y);
        t = ff_add1<COMP>(t, x[k].z);
        t = ff_add1<COMP>(t, x[k].w);
        const ff s = wave_incl_scan_ff<COMP>(t);
        // exclusive = inclusive of the lane below (lane 0: zero)
        ff e;
        e.s = GPS_DPP_F(s.s, 0x138, 0xf, true);  // wave_shr:1
        e.c = COMP ? GPS_DPP_F(s.c, 0x138, 0xf, true) : 0.f;
        lane_carry[k] = ff_add<COMP>(run, e);
        const ff tot = {readlane_f(s.s, 63), COMP ? readlane_f(s.c, 63) : 0.f};
        run = ff_add<COMP>(run, tot);
    }
    if (lane == 0) {
        s_ws[wave] = run.s;
        s_wc[wave] = run.c;
    }
    __syncthreads();
    // Wave totals: lane q takes wave q's total and one pair wave scan gives
    // every wave its exclusive prefix and the tile total (a serial loop over
    // the totals held 32 more registers: 102 against 79 VGPRs at 1024 x 8).
    ff wave_pre, tile_tot;
    {
        const uint32_t q = lane < C::kWaves ? lane : 0u;
        const ff v = {lane < C::kWaves ? s_ws[q] : 0.f, lane < C::kWaves ? s_wc[q] : 0.f};
        const ff sc = wave_incl_scan_ff<COMP>(v);
        const ff ex = {GPS_DPP_F(sc.s, 0x138, 0xf, true), COMP ? GPS_DPP_F(sc.c, 0x138, 0xf, true) : 0.f};
        wave_pre = {readlane_f(ex.s, (int)wave), COMP ? readlane_f(ex.c, (int)wave) : 0.f};
        tile_tot = {readlane_f(sc.s, C::kWaves - 1), COMP ? readlane_f(sc.c, C::kWaves - 1) : 0.f};
    }
    if (wave == 0) {
        ff prefix = {0.f, 0.f};
        if (tile == 0) {
            if (lane == 0) publish_ff<COMP>(p.status, 0, tile_tot, kFlagIncl, epoch);
        } else {
            if (lane == 0) publish_ff<COMP>(p.status, tile, tile_tot, kFlagAgg, epoch);
            prefix = lookback_ff<COMP>(p.status, tile, epoch, p.err);
            if (lane == 0) publish_ff<COMP>(p.status, tile, ff_add<COMP>(prefix, tile_tot), kFlagIncl, epoch);
        }
        if (lane == 0) {
            s_ps = prefix.s;
            s_pc = prefix.c;
        }
    }
    __syncthreads();
    const ff base_carry = ff_add<COMP>({s_ps, s_pc}, wave_pre);
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        ff acc = ff_add<COMP>(base_carry, lane_carry[k]);
        f32x4 y;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            acc = ff_add1<COMP>(acc, x[k][c]);
            y[c] = acc.s + acc.c;
        }
        x[k] = y;
    }
    if (vec_path) {
#pragma unroll
        for (int k = 0; k < VEC; ++k)
            __builtin_nontemporal_store(x[k], reinterpret_cast<f32x4*>(p.out + base + (uint64_t)k * 256u));
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
                if (i < p.n) p.out[i] = x[k][c];
            }
    }
}

// ---------------------------------------------------------------------------
// Persistent, software-pipelined FP scan (the uint32 csdl_pp_u32 schedule):
// resident workgroups loop over ticketed tiles; with tile i scanned and its
// aggregate published, a workgroup takes its next ticket and issues tile
// i+1's loads, then resolves tile i's lookback and stores tile i. The carry
// arithmetic comes from a policy, PairCarry<COMP> (pair or plain floats),
// with the same per-level operations and order as scan_fp.
//
// 2^27 FP32 / BF16 elements at 1024 x 8: pairs 4998 / 4433 GB/s against
// 4787 / 4009 for one workgroup per tile, plain 5567 / 5285 against 5383 /
// 4921; at 2^23 one workgroup per tile is faster.
// ---------------------------------------------------------------------------
template <bool COMP>
struct PairCarry {
    using T = ff;
    static __device__ __forceinline__ T zero() { return {0.f, 0.f}; }
    static __device__ __forceinline__ T add(T a, T b) { return ff_add<COMP>(a, b); }
    static __device__ __forceinline__ T add1(T a, float v) { return ff_add1<COMP>(a, v); }
    static __device__ __forceinline__ T vec_total(const f32x4& v) {
        T t = {v.x, 0.f};
        t = ff_add1<COMP>(t, v.y);
        t = ff_add1<COMP>(t, v.z);
        return ff_add1<COMP>(t, v.w);
    }
    static __device__ __forceinline__ T scan(T x) { return wave_incl_scan_ff<COMP>(x); }
    static __device__ __forceinline__ T shr1(T x) {  // wave_shr:1, lane 0 gets zero
        return {GPS_DPP_F(x.s, 0x138, 0xf, true), COMP ? GPS_DPP_F(x.c, 0x138, 0xf, true) : 0.f};
    }
    static __device__ __forceinline__ T lane(T x, int l) { return {readlane_f(x.s, l), COMP ? readlane_f(x.c, l) : 0.f}; }
    static __device__ __forceinline__ void publish(uint64_t* st, uint32_t tile, T v, uint64_t flag, uint32_t epoch) {
        publish_ff<COMP>(st, tile, v, flag, epoch);
    }
    static __device__ __forceinline__ T from_words(uint64_t w0, uint64_t w1) {
        return {__builtin_bit_cast(float, (uint32_t)w0), __builtin_bit_cast(float, (uint32_t)w1)};
    }
    static __device__ __forceinline__ float round(T a) { return a.s + a.c; }
};

// Window lookback over two-word tile values, split into issue() and consume()
// so the first poll can be issued ahead of the next tile's loads.
template <class Ops>
struct LookbackW2 {
    typename Ops::T prefix;
    int32_t top;
    uint64_t w0, w1;
    uint32_t empty_polls;

    __device__ __forceinline__ void start(uint32_t tile) {
        prefix = Ops::zero();
        top = (int32_t)tile - 1;
        empty_polls = 0;
    }
    __device__ __forceinline__ void issue(const uint64_t* status, uint32_t epoch) {
        const int32_t idx = top - (int32_t)lane_id();
        const uint64_t at = (uint64_t)(idx >= 0 ? idx : 0);
        uint64_t l0, l1;
        load_pair_words(status + kSlotWords * at, l0, l1);
        const uint64_t sentinel = make_status(0u, kFlagIncl, epoch);
        w0 = idx >= 0 ? l0 : sentinel;
        w1 = idx >= 0 ? l1 : sentinel;
    }
    // True once `prefix` is complete (or the watchdog fired).
    __device__ __forceinline__ bool consume(uint32_t epoch, uint32_t* err) {
        const uint32_t lane = lane_id();
        const uint32_t f0 = status_flag(w0, epoch), f1 = status_flag(w1, epoch);
        const uint32_t f = f0 == f1 ? f0 : 0u;
        const uint64_t ready = ballot(f != 0u);
        const uint64_t incl = ballot(f == (uint32_t)kFlagIncl);
        const uint32_t run = (~ready) ? (uint32_t)__builtin_ctzll(~ready) : 64u;
        const uint32_t first_incl = incl ? (uint32_t)__builtin_ctzll(incl) : 64u;
        const uint32_t take = first_incl < run ? first_incl + 1u : run;
        if (take) {
            const typename Ops::T v = lane < take ? Ops::from_words(w0, w1) : Ops::zero();
            prefix = Ops::add(Ops::lane(Ops::scan(v), 63), prefix);
            if (first_incl < run) return true;
            top -= (int32_t)run;
            empty_polls = 0;
        } else {
            if (watchdog_expired(empty_polls, err)) {
                if (lane == 0) raise_device_error(err, kErrLookbackWatchdog);
                return true;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return false;
    }
};

template <int BLOCK, int VEC, bool BF16, class Ops>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void scan_fp_pp(
    FpParams p) {
    using C = TileCfg<BLOCK, VEC>;
    using T = typename Ops::T;
    __shared__ T s_w[C::kWaves];
    __shared__ T s_p;
    __shared__ uint32_t s_ticket[3];
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

    f32x4 x[VEC], y[VEC];
    load_fp<BF16, VEC>(x, p.in, thread_base<BLOCK, VEC>(tile), p.n, p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n));

    while (true) {
        // Per vector: lane total, wave exclusive scan -> lane carry.
        T lane_carry[VEC];
        T run = Ops::zero();
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const T s = Ops::scan(Ops::vec_total(x[k]));
            lane_carry[k] = Ops::add(run, Ops::shr1(s));
            run = Ops::add(run, Ops::lane(s, 63));
        }
        if (lane == 0) s_w[wave] = run;
        __syncthreads();
        T wave_pre, tile_tot;
        {
            const T v = lane < C::kWaves ? s_w[lane] : Ops::zero();
            const T sc = Ops::scan(v);
            wave_pre = Ops::lane(Ops::shr1(sc), (int)wave);
            tile_tot = Ops::lane(sc, C::kWaves - 1);
        }
        if (threadIdx.x == 0) {
            Ops::publish(p.status, tile, tile_tot, tile ? kFlagAgg : kFlagIncl, epoch);
            s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
        }
        __syncthreads();
        const uint32_t next = s_ticket[2];
        const bool has_next = next < p.num_tiles;

        LookbackW2<Ops> lb;
        if (wave == 0 && tile != 0) {
            lb.start(tile);
            lb.issue(p.status, epoch);
        }
        if (has_next)
            load_fp<BF16, VEC>(y, p.in, thread_base<BLOCK, VEC>(next), p.n,
                               p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n));
        if (wave == 0) {
            T prefix = Ops::zero();
            if (tile != 0) {
                while (!lb.consume(epoch, p.err)) lb.issue(p.status, epoch);
                prefix = lb.prefix;
                if (lane == 0) Ops::publish(p.status, tile, Ops::add(prefix, tile_tot), kFlagIncl, epoch);
            }
            if (lane == 0) s_p = prefix;
        }
        __syncthreads();
        const T base_carry = Ops::add(s_p, wave_pre);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            T acc = Ops::add(base_carry, lane_carry[k]);
            f32x4 o;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                acc = Ops::add1(acc, x[k][c]);
                o[c] = Ops::round(acc);
            }
            x[k] = o;
        }
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n)) {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
                __builtin_nontemporal_store(x[k], reinterpret_cast<f32x4*>(p.out + base + (uint64_t)k * 256u));
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
                    if (i < p.n) p.out[i] = x[k][c];
                }
        }
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) x[k] = y[k];
        tile = next;
    }
}

// ---------------------------------------------------------------------------
// Batched segmented FP inclusive scan: one launch over concatenated segments
// (segment s covers [offsets[s], offsets[s+1])), heads found as in the uint32
// segmented scan (segments.hpp). A floating-point carry cannot be cut at a
// head by subtracting prefixes (cancellation), so every level applies the
// segmented operator to (flag, pair):
//     (f1, v1) + (f2, v2) = (f1 | f2, f2 ? v2 : v1 + v2)
// in the lane, in the DPP wave scan, across vectors, across waves and along
// the lookback chain. A tile that contains a head publishes INCLUSIVE at once
// (its carry-out is the sum since its last head); only elements before the
// tile's first head take the looked-back prefix.
// ---------------------------------------------------------------------------
#define GPS_DPP_I(x, CTRL, RM, BC) ((uint32_t)__builtin_amdgcn_update_dpp(0, (int)(x), CTRL, RM, 0xf, BC))

// Inclusive segmented wave scan of (f, x). Lanes that a DPP step does not
// feed (bound_ctrl, or rows outside the broadcast mask) receive (0, 0), the
// identity, so the same six-step schedule as the plain scan applies.
template <bool COMP>
__device__ __forceinline__ void wave_seg_scan_ff(uint32_t& f, ff& x) {
#define GPS_SEG_STEP(CTRL, RM, BC)                                                                  \
    {                                                                                               \
        const ff y = {GPS_DPP_F(x.s, CTRL, RM, BC), COMP ? GPS_DPP_F(x.c, CTRL, RM, BC) : 0.f};     \
        const uint32_t yf = GPS_DPP_I(f, CTRL, RM, BC);                                             \
        const ff sum = ff_add<COMP>(y, x);                                                          \
        x = f ? x : sum;                                                                            \
        f |= yf;                                                                                    \
    }
    GPS_SEG_STEP(0x111, 0xf, true)
    GPS_SEG_STEP(0x112, 0xf, true)
    GPS_SEG_STEP(0x114, 0xf, true)
    GPS_SEG_STEP(0x118, 0xf, true)
    GPS_SEG_STEP(0x142, 0xa, false)
    GPS_SEG_STEP(0x143, 0xc, false)
#undef GPS_SEG_STEP
}

// acc <- acc (+) (f, v) under the segmented operator (acc on the left).
template <bool COMP>
__device__ __forceinline__ void seg_combine_ff(uint32_t& acc_f, ff& acc, uint32_t f, ff v) {
    acc = f ? v : ff_add<COMP>(acc, v);
    acc_f |= f;
}

struct SegFpParams {
    const void* in;
    float* out;
    const uint64_t* offsets;
    uint64_t num_segments;
    uint64_t n;
    uint32_t num_tiles;
    uint32_t vec_ok;
    uint64_t* status;  // 2 words per tile
    uint64_t* ticket;
    uint32_t* err;
};

template <int BLOCK, int VEC, bool BF16, bool COMP>
__global__ __launch_bounds__(BLOCK) void segscan_fp(SegFpParams p) {
    using C = TileCfg<BLOCK, VEC>;
    constexpr uint32_t kWords = C::kTile / 32;
    __shared__ uint32_t s_heads[kWords];
    __shared__ float s_ws[C::kWaves], s_wc[C::kWaves];
    __shared__ uint32_t s_wf[C::kWaves];
    __shared__ uint32_t s_ticket[2];
    __shared__ float s_ps, s_pc;
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
    const uint64_t base = thread_base<BLOCK, VEC>(tile);
    const bool vec_path = p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n);
    f32x4 x[VEC];
    load_fp<BF16, VEC>(x, p.in, base, p.n, vec_path);  // the head search overlaps these loads

    if (wave == 0) {
        const uint64_t s0 = find_first_segment(p.offsets, p.num_segments, a);
        if (lane == 0) s_first_seg = s0;
    }
    __syncthreads();
    mark_heads<BLOCK>(s_heads, p.offsets, p.num_segments, s_first_seg, a, b, tile == 0);
    __syncthreads();

    // Per vector: lane-local segmented sums, the segmented wave scan, and the
    // lane's carry-in. `open` bit k: no head precedes vector k's lane in the
    // wave's run, so the carry from earlier waves / tiles still applies.
    const uint32_t run_base = wave * C::kWaveElems;
    uint32_t nibs[VEC];
    ff lane_carry[VEC];
    uint32_t open = 0;
    ff run = {0.f, 0.f};
    uint32_t run_f = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const uint32_t nib = run_nibble(s_heads + run_base / 32u + lane / 8u, k, lane);
        nibs[k] = nib;
        ff t = {x[k].x, 0.f};
        t = (nib & 2u) ? ff{x[k].y, 0.f} : ff_add1<COMP>(t, x[k].y);
        t = (nib & 4u) ? ff{x[k].z, 0.f} : ff_add1<COMP>(t, x[k].z);
        t = (nib & 8u) ? ff{x[k].w, 0.f} : ff_add1<COMP>(t, x[k].w);
        // A vector with no head in any lane (the common case for long
        // segments) takes the plain pair scan: no flag moves, no selects.
        const bool any_head = ballot(nib != 0u) != 0ull;
        uint32_t sf = 0u, ef = 0u;
        ff sv = t;
        if (any_head) {
            sf = nib != 0u;
            wave_seg_scan_ff<COMP>(sf, sv);
            ef = GPS_DPP_I(sf, 0x138, 0xf, true);  // wave_shr:1, lane 0 gets 0
        } else {
            sv = wave_incl_scan_ff<COMP>(t);
        }
        const ff ev = {GPS_DPP_F(sv.s, 0x138, 0xf, true), COMP ? GPS_DPP_F(sv.c, 0x138, 0xf, true) : 0.f};
        lane_carry[k] = ef ? ev : ff_add<COMP>(run, ev);
        open |= (ef | run_f) ? 0u : (1u << k);
        const uint32_t tf = readlane(sf, 63);
        const ff tv = {readlane_f(sv.s, 63), COMP ? readlane_f(sv.c, 63) : 0.f};
        seg_combine_ff<COMP>(run_f, run, tf, tv);
    }
    if (lane == 0) {
        s_ws[wave] = run.s;
        s_wc[wave] = run.c;
        s_wf[wave] = run_f;
    }
    __syncthreads();
    // Wave totals: lane q takes wave q's (flag, pair) and one segmented wave
    // scan gives every wave its exclusive prefix and the tile total.
    uint32_t pre_f, tile_f;
    ff pre, tile_tot;
    {
        const uint32_t q = lane < C::kWaves ? lane : 0u;
        uint32_t f = lane < C::kWaves ? s_wf[q] : 0u;
        ff v = {lane < C::kWaves ? s_ws[q] : 0.f, lane < C::kWaves ? s_wc[q] : 0.f};
        wave_seg_scan_ff<COMP>(f, v);
        const uint32_t ef = GPS_DPP_I(f, 0x138, 0xf, true);  // wave_shr:1, lane 0 gets (0, 0)
        const ff ev = {GPS_DPP_F(v.s, 0x138, 0xf, true), COMP ? GPS_DPP_F(v.c, 0x138, 0xf, true) : 0.f};
        pre_f = readlane(ef, (int)wave);
        pre = {readlane_f(ev.s, (int)wave), COMP ? readlane_f(ev.c, (int)wave) : 0.f};
        tile_f = readlane(f, C::kWaves - 1);
        tile_tot = {readlane_f(v.s, C::kWaves - 1), COMP ? readlane_f(v.c, C::kWaves - 1) : 0.f};
    }
    const bool starts_with_head = (s_heads[0] & 1u) != 0u;
    if (wave == 0) {
        if (lane == 0)
            publish_ff<COMP>(p.status, tile, tile_tot, (tile_f || tile == 0) ? kFlagIncl : kFlagAgg, epoch);
        ff prefix = {0.f, 0.f};
        if (tile != 0 && !starts_with_head) {
            prefix = lookback_ff<COMP>(p.status, tile, epoch, p.err);
            if (!tile_f && lane == 0)
                publish_ff<COMP>(p.status, tile, ff_add<COMP>(prefix, tile_tot), kFlagIncl, epoch);
        }
        if (lane == 0) {
            s_ps = prefix.s;
            s_pc = prefix.c;
        }
    }
    __syncthreads();
    const ff run_in = pre_f ? pre : ff_add<COMP>({s_ps, s_pc}, pre);
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        ff acc = (open >> k) & 1u ? ff_add<COMP>(run_in, lane_carry[k]) : lane_carry[k];
        f32x4 y;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if ((nibs[k] >> c) & 1u) acc = {0.f, 0.f};
            acc = ff_add1<COMP>(acc, x[k][c]);
            y[c] = acc.s + acc.c;
        }
        x[k] = y;
    }
    if (vec_path) {
#pragma unroll
        for (int k = 0; k < VEC; ++k)
            __builtin_nontemporal_store(x[k], reinterpret_cast<f32x4*>(p.out + base + (uint64_t)k * 256u));
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
                if (i < p.n) p.out[i] = x[k][c];
            }
    }
}

// Persistent, software-pipelined form of segscan_fp (the segscan_pp_u32
// schedule): while wave 0 resolves tile i's lookback, the last wave finds and
// marks tile i+1's heads into the other half of a double-buffered LDS bitmap
// and every wave issues tile i+1's loads. The output pass re-reads tile i's
// head nibbles from its bitmap instead of holding them in registers beside
// the two register tiles.
template <int BLOCK, int VEC, bool BF16, bool COMP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void segscan_fp_pp(SegFpParams p) {
    using C = TileCfg<BLOCK, VEC>;
    using Ops = PairCarry<COMP>;
    constexpr uint32_t kWords = C::kTile / 32;
    constexpr uint32_t kSearchWave = C::kWaves > 1 ? C::kWaves - 1 : 0;
    __shared__ uint32_t s_heads[2][kWords];
    __shared__ float s_ws[C::kWaves], s_wc[C::kWaves];
    __shared__ uint32_t s_wf[C::kWaves];
    __shared__ uint32_t s_ticket[3];
    __shared__ float s_ps, s_pc;
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

    f32x4 x[VEC], y[VEC];
    load_fp<BF16, VEC>(x, p.in, thread_base<BLOCK, VEC>(tile), p.n, p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n));
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

    const uint32_t run_base = wave * C::kWaveElems;
    uint32_t cur = 0;
    while (true) {
        // Scan phase, as segscan_fp: lane-local segmented sums, the segmented
        // wave scan and each lane's carry-in per vector.
        const uint32_t* lane_words = s_heads[cur] + run_base / 32u + lane / 8u;
        ff lane_carry[VEC];
        uint32_t open = 0;
        ff run = {0.f, 0.f};
        uint32_t run_f = 0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const uint32_t nib = run_nibble(lane_words, k, lane);
            ff t = {x[k].x, 0.f};
            t = (nib & 2u) ? ff{x[k].y, 0.f} : ff_add1<COMP>(t, x[k].y);
            t = (nib & 4u) ? ff{x[k].z, 0.f} : ff_add1<COMP>(t, x[k].z);
            t = (nib & 8u) ? ff{x[k].w, 0.f} : ff_add1<COMP>(t, x[k].w);
            const bool any_head = ballot(nib != 0u) != 0ull;
            uint32_t sf = 0u, ef = 0u;
            ff sv = t;
            if (any_head) {
                sf = nib != 0u;
                wave_seg_scan_ff<COMP>(sf, sv);
                ef = GPS_DPP_I(sf, 0x138, 0xf, true);  // wave_shr:1, lane 0 gets 0
            } else {
                sv = wave_incl_scan_ff<COMP>(t);
            }
            const ff ev = Ops::shr1(sv);
            lane_carry[k] = ef ? ev : ff_add<COMP>(run, ev);
            open |= (ef | run_f) ? 0u : (1u << k);
            const uint32_t tf = readlane(sf, 63);
            seg_combine_ff<COMP>(run_f, run, tf, Ops::lane(sv, 63));
        }
        if (lane == 0) {
            s_ws[wave] = run.s;
            s_wc[wave] = run.c;
            s_wf[wave] = run_f;
        }
        __syncthreads();
        uint32_t pre_f, tile_f;
        ff pre, tile_tot;
        {
            const uint32_t q = lane < C::kWaves ? lane : 0u;
            uint32_t f = lane < C::kWaves ? s_wf[q] : 0u;
            ff v = {lane < C::kWaves ? s_ws[q] : 0.f, lane < C::kWaves ? s_wc[q] : 0.f};
            wave_seg_scan_ff<COMP>(f, v);
            const uint32_t ef = GPS_DPP_I(f, 0x138, 0xf, true);
            pre_f = readlane(ef, (int)wave);
            pre = Ops::lane(Ops::shr1(v), (int)wave);
            tile_f = readlane(f, C::kWaves - 1);
            tile_tot = Ops::lane(v, C::kWaves - 1);
        }
        const bool need_lookback = tile != 0 && (s_heads[cur][0] & 1u) == 0u;
        if (threadIdx.x == 0) {
            // a tile with a head knows its carry-out: publish it inclusive now
            publish_ff<COMP>(p.status, tile, tile_tot, (tile_f || tile == 0) ? kFlagIncl : kFlagAgg, epoch);
            s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
        }
        // the other head buffer was last read by the previous tile's output pass
        for (uint32_t i = threadIdx.x; i < kWords; i += BLOCK) s_heads[cur ^ 1u][i] = 0u;
        __syncthreads();
        const uint32_t next = s_ticket[2];
        const bool has_next = next < p.num_tiles;

        LookbackW2<Ops> lb;
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
        if (has_next)
            load_fp<BF16, VEC>(y, p.in, thread_base<BLOCK, VEC>(next), p.n,
                               p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n));
        if (wave == 0) {
            ff prefix = {0.f, 0.f};
            if (need_lookback) {
                while (!lb.consume(epoch, p.err)) lb.issue(p.status, epoch);
                prefix = lb.prefix;
                if (!tile_f && lane == 0)
                    publish_ff<COMP>(p.status, tile, ff_add<COMP>(prefix, tile_tot), kFlagIncl, epoch);
            }
            if (lane == 0) {
                s_ps = prefix.s;
                s_pc = prefix.c;
            }
        }
        __syncthreads();  // also completes the next tile's head bitmap
        const ff run_in = pre_f ? pre : ff_add<COMP>({s_ps, s_pc}, pre);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const uint32_t nib = run_nibble(lane_words, k, lane);
            ff acc = (open >> k) & 1u ? ff_add<COMP>(run_in, lane_carry[k]) : lane_carry[k];
            f32x4 o;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if ((nib >> c) & 1u) acc = {0.f, 0.f};
                acc = ff_add1<COMP>(acc, x[k][c]);
                o[c] = acc.s + acc.c;
            }
            x[k] = o;
        }
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n)) {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
                __builtin_nontemporal_store(x[k], reinterpret_cast<f32x4*>(p.out + base + (uint64_t)k * 256u));
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
                    if (i < p.n) p.out[i] = x[k][c];
                }
        }
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) x[k] = y[k];
        tile = next;
        cur ^= 1u;
    }
}

// ---------------------------------------------------------------------------
// Persistent FP scan over lane-blocked runs (GPS_SCHEDULE_PERSISTENT with
// pair carries; the default above 2^24 elements).
//
// scan_fp_pp pays one pair wave scan per uint4 vector: VEC scans of 64 lane
// totals per wave run, about 60 VALU instructions each, most of the pair
// path's ~38 VALU instructions per element. Here each wave transposes its
// run through its own slice of LDS after the coalesced loads, so lane l holds
// the 4*VEC consecutive elements 4*VEC*l .. 4*VEC*l + 4*VEC - 1; the lane
// sums them serially (TwoSum pairs), one pair wave scan per run gives every
// lane its carry, the outputs are recomputed serially from that carry, and
// the outputs are transposed back for coalesced stores. Per element that is
// one TwoSum add for the total and one for the output, plus 1/(4*VEC) of a
// wave scan; the two transposes are four LDS instructions per lane-vector.
//
// LDS slice per wave: 64 rows (one per lane) of 4*VEC floats padded by 4, so
// that eight consecutive lanes' 16-byte accesses, row-wise or column-wise,
// fall in distinct bank groups. Each wave reads only what it wrote; LDS
// instructions of one wave complete in order, and a wavefront-scope fence
// keeps the compiler from reordering the cross-lane write/read pairs.
// ---------------------------------------------------------------------------
template <int VEC>
struct BlockedSlice {
    static constexpr uint32_t kRow = 4u * VEC + 4u;  // floats per padded row
    static constexpr uint32_t kFloats = 64u * kRow;
};

// Striped (lane l, vector k) <-> blocked (lane l, vector q) through the
// wave's LDS slice. Striped element of (k, l, c): 256k + 4l + c; blocked
// element of (l, q, c): 4*VEC*l + 4q + c.
template <int VEC>
__device__ __forceinline__ void striped_to_blocked(f32x4 (&x)[VEC], float* slice) {
    using S = BlockedSlice<VEC>;
    const uint32_t l = lane_id();
    constexpr uint32_t kRowsPerVec = 256u / (4u * VEC);  // rows covered by one striped vector
    constexpr uint32_t kLanesPerRow = VEC;                // striped lanes that share a row
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const uint32_t row = k * kRowsPerVec + l / kLanesPerRow;
        *reinterpret_cast<f32x4*>(slice + row * S::kRow + 4u * (l % kLanesPerRow)) = x[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < VEC; ++q) x[q] = *reinterpret_cast<const f32x4*>(slice + l * S::kRow + 4u * q);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int VEC>
__device__ __forceinline__ void blocked_to_striped(f32x4 (&x)[VEC], float* slice) {
    using S = BlockedSlice<VEC>;
    const uint32_t l = lane_id();
    constexpr uint32_t kRowsPerVec = 256u / (4u * VEC);
    constexpr uint32_t kLanesPerRow = VEC;
#pragma unroll
    for (int q = 0; q < VEC; ++q) *reinterpret_cast<f32x4*>(slice + l * S::kRow + 4u * q) = x[q];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const uint32_t row = k * kRowsPerVec + l / kLanesPerRow;
        x[k] = *reinterpret_cast<const f32x4*>(slice + row * S::kRow + 4u * (l % kLanesPerRow));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int BLOCK, int VEC, bool BF16, class Ops>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void scan_fp_bl(FpParams p) {
    using C = TileCfg<BLOCK, VEC>;
    using T = typename Ops::T;
    static_assert(VEC == 2 || VEC == 4 || VEC == 8, "blocked runs need 256 / (4 VEC) whole rows per vector");
    __shared__ float s_slices[C::kWaves * BlockedSlice<VEC>::kFloats];
    __shared__ T s_w[C::kWaves];
    __shared__ T s_p;
    __shared__ uint32_t s_ticket[3];
    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    const uint32_t total_tickets = p.num_tiles + gridDim.x;
    float* slice = s_slices + wave * BlockedSlice<VEC>::kFloats;

    if (threadIdx.x == 0) {
        const Ticket t = take_ticket(p.ticket, total_tickets);
        s_ticket[0] = t.tile;
        s_ticket[1] = t.epoch;
    }
    __syncthreads();
    uint32_t tile = s_ticket[0];
    const uint32_t epoch = s_ticket[1];
    if (tile >= p.num_tiles) return;

    f32x4 x[VEC], y[VEC];
    load_fp<BF16, VEC>(x, p.in, thread_base<BLOCK, VEC>(tile), p.n, p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n));

    while (true) {
        // Lane total of the lane's consecutive run, then one pair wave scan.
        striped_to_blocked<VEC>(x, slice);
        T t = {x[0].x, 0.f};
        t = Ops::add1(t, x[0].y);
        t = Ops::add1(t, x[0].z);
        t = Ops::add1(t, x[0].w);
#pragma unroll
        for (int q = 1; q < VEC; ++q) {
#pragma unroll
            for (int c = 0; c < 4; ++c) t = Ops::add1(t, x[q][c]);
        }
        const T sc = Ops::scan(t);
        const T lane_carry = Ops::shr1(sc);
        if (lane == 0) s_w[wave] = Ops::lane(sc, 63);
        __syncthreads();
        T wave_pre, tile_tot;
        {
            const T v = lane < C::kWaves ? s_w[lane] : Ops::zero();
            const T ws = Ops::scan(v);
            wave_pre = Ops::lane(Ops::shr1(ws), (int)wave);
            tile_tot = Ops::lane(ws, C::kWaves - 1);
        }
        if (threadIdx.x == 0) {
            Ops::publish(p.status, tile, tile_tot, tile ? kFlagAgg : kFlagIncl, epoch);
            s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
        }
        __syncthreads();
        const uint32_t next = s_ticket[2];
        const bool has_next = next < p.num_tiles;

        LookbackW2<Ops> lb;
        if (wave == 0 && tile != 0) {
            lb.start(tile);
            lb.issue(p.status, epoch);
        }
        if (has_next)
            load_fp<BF16, VEC>(y, p.in, thread_base<BLOCK, VEC>(next), p.n,
                               p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n));
        if (wave == 0) {
            T prefix = Ops::zero();
            if (tile != 0) {
                while (!lb.consume(epoch, p.err)) lb.issue(p.status, epoch);
                prefix = lb.prefix;
                if (lane == 0) Ops::publish(p.status, tile, Ops::add(prefix, tile_tot), kFlagIncl, epoch);
            }
            if (lane == 0) s_p = prefix;
        }
        __syncthreads();
        // Outputs: the lane's carry, then its run serially, each rounded once.
        T acc = Ops::add(Ops::add(s_p, wave_pre), lane_carry);
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
            f32x4 o;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                acc = Ops::add1(acc, x[q][c]);
                o[c] = Ops::round(acc);
            }
            x[q] = o;
        }
        blocked_to_striped<VEC>(x, slice);
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n)) {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
                __builtin_nontemporal_store(x[k], reinterpret_cast<f32x4*>(p.out + base + (uint64_t)k * 256u));
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
                    if (i < p.n) p.out[i] = x[k][c];
                }
        }
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) x[k] = y[k];
        tile = next;
    }
}

// ---------------------------------------------------------------------------
// Persistent segmented FP scan over lane-blocked runs (pair carries; the
// persistent schedule's kernel for them). segscan_fp_pp's tile loop and head
// discovery with scan_fp_bl's per-wave LDS transpose: lane l of wave w holds
// tile elements w*256*VEC + 4*VEC*l .. + 4*VEC - 1, whose head bits are
// 4*VEC consecutive bits of the tile's head bitmap (one word at VEC = 8). The
// lane scans its run serially with the carry reset at heads, one segmented
// pair wave scan per run gives every lane its carry-in, and the outputs are
// recomputed serially and transposed back.
// ---------------------------------------------------------------------------
template <int VEC>
__device__ __forceinline__ uint32_t lane_run_heads(const uint32_t* s_heads, uint32_t run_base, uint32_t lane) {
    constexpr uint32_t E = 4u * VEC;  // elements per lane run
    const uint32_t first = run_base + E * lane;
    const uint32_t word = s_heads[first / 32u];
    if (E == 32u) return word;
    return (word >> (first % 32u)) & ((1u << E) - 1u);
}

template <int BLOCK, int VEC, bool BF16, bool COMP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void segscan_fp_bl(SegFpParams p) {
    using C = TileCfg<BLOCK, VEC>;
    using Ops = PairCarry<COMP>;
    constexpr uint32_t kWords = C::kTile / 32;
    constexpr uint32_t kSearchWave = C::kWaves > 1 ? C::kWaves - 1 : 0;
    __shared__ float s_slices[C::kWaves * BlockedSlice<VEC>::kFloats];
    __shared__ uint32_t s_heads[2][kWords];
    __shared__ float s_ws[C::kWaves], s_wc[C::kWaves];
    __shared__ uint32_t s_wf[C::kWaves];
    __shared__ uint32_t s_ticket[3];
    __shared__ float s_ps, s_pc;
    __shared__ uint64_t s_first_seg;
    const uint32_t wave = wave_index();
    const uint32_t lane = lane_id();
    const uint32_t total_tickets = p.num_tiles + gridDim.x;
    float* slice = s_slices + wave * BlockedSlice<VEC>::kFloats;
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

    f32x4 x[VEC], y[VEC];
    load_fp<BF16, VEC>(x, p.in, thread_base<BLOCK, VEC>(tile), p.n, p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n));
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

    const uint32_t run_base = wave * C::kWaveElems;
    uint32_t cur = 0;
    while (true) {
        // Lane run: segmented serial sum (carry reset at heads), then one
        // segmented pair wave scan for the lane's carry-in.
        striped_to_blocked<VEC>(x, slice);
        const uint32_t heads = lane_run_heads<VEC>(s_heads[cur], run_base, lane);
        // The run's segmented total is the sum from its last head on: the
        // elements before that head add 0 to the zero pair (exact), so one
        // select on the input replaces a reset of the pair at every head;
        // the sum is bit-identical. (At C3: 0.086-0.088 -> 0.081-0.083 ms.)
        ff t = {0.f, 0.f};
        {
            const uint32_t last = heads ? 31u - (uint32_t)__builtin_clz(heads) : 0u;
#pragma unroll
            for (int q = 0; q < VEC; ++q) {
#pragma unroll
                for (int c = 0; c < 4; ++c) t = Ops::add1(t, (uint32_t)(4 * q + c) >= last ? x[q][c] : 0.f);
            }
        }
        uint32_t sf = heads != 0u;
        ff sv = t;
        wave_seg_scan_ff<COMP>(sf, sv);
        const uint32_t ef = GPS_DPP_I(sf, 0x138, 0xf, true);  // wave_shr:1, lane 0 gets 0
        const ff ev = Ops::shr1(sv);
        if (lane == 0) {
            s_ws[wave] = readlane_f(sv.s, 63);
            s_wc[wave] = COMP ? readlane_f(sv.c, 63) : 0.f;
            s_wf[wave] = readlane(sf, 63);
        }
        __syncthreads();
        uint32_t pre_f, tile_f;
        ff pre, tile_tot;
        {
            const uint32_t q = lane < C::kWaves ? lane : 0u;
            uint32_t f = lane < C::kWaves ? s_wf[q] : 0u;
            ff v = {lane < C::kWaves ? s_ws[q] : 0.f, lane < C::kWaves ? s_wc[q] : 0.f};
            wave_seg_scan_ff<COMP>(f, v);
            const uint32_t wef = GPS_DPP_I(f, 0x138, 0xf, true);
            pre_f = readlane(wef, (int)wave);
            pre = Ops::lane(Ops::shr1(v), (int)wave);
            tile_f = readlane(f, C::kWaves - 1);
            tile_tot = Ops::lane(v, C::kWaves - 1);
        }
        const bool need_lookback = tile != 0 && (s_heads[cur][0] & 1u) == 0u;
        if (threadIdx.x == 0) {
            // a tile with a head knows its carry-out: publish it inclusive now
            publish_ff<COMP>(p.status, tile, tile_tot, (tile_f || tile == 0) ? kFlagIncl : kFlagAgg, epoch);
            s_ticket[2] = take_ticket(p.ticket, total_tickets).tile;
        }
        // the other head buffer was last read by the previous tile's output pass
        for (uint32_t i = threadIdx.x; i < kWords; i += BLOCK) s_heads[cur ^ 1u][i] = 0u;
        __syncthreads();
        const uint32_t next = s_ticket[2];
        const bool has_next = next < p.num_tiles;

        LookbackW2<Ops> lb;
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
        if (has_next)
            load_fp<BF16, VEC>(y, p.in, thread_base<BLOCK, VEC>(next), p.n,
                               p.vec_ok && tile_is_full<BLOCK, VEC>(next, p.n));
        if (wave == 0) {
            ff prefix = {0.f, 0.f};
            if (need_lookback) {
                while (!lb.consume(epoch, p.err)) lb.issue(p.status, epoch);
                prefix = lb.prefix;
                if (!tile_f && lane == 0)
                    publish_ff<COMP>(p.status, tile, ff_add<COMP>(prefix, tile_tot), kFlagIncl, epoch);
            }
            if (lane == 0) {
                s_ps = prefix.s;
                s_pc = prefix.c;
            }
        }
        __syncthreads();  // also completes the next tile's head bitmap
        // The lane's carry-in: its segmented exclusive value, plus the run's
        // carry-in when no head precedes the lane in its wave.
        const ff run_in = pre_f ? pre : ff_add<COMP>({s_ps, s_pc}, pre);
        ff acc = ef ? ev : ff_add<COMP>(run_in, ev);
        // Lanes whose run holds a head reset the pair at each one; the rest
        // (nearly all, at thousands of elements per segment) skip the resets.
        // The two branches write x in place, so they share its registers
        // (128 VGPRs at 1024 x 8, no spill); a wave with no head lane skips
        // the first. (At C3: 0.081-0.082 -> 0.080-0.081 ms.)
        if (heads != 0u) {
#pragma unroll
            for (int q = 0; q < VEC; ++q) {
                f32x4 o;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if ((heads >> (4 * q + c)) & 1u) acc = {0.f, 0.f};
                    acc = Ops::add1(acc, x[q][c]);
                    o[c] = Ops::round(acc);
                }
                x[q] = o;
            }
        } else {
#pragma unroll
            for (int q = 0; q < VEC; ++q) {
                f32x4 o;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc = Ops::add1(acc, x[q][c]);
                    o[c] = Ops::round(acc);
                }
                x[q] = o;
            }
        }
        blocked_to_striped<VEC>(x, slice);
        const uint64_t base = thread_base<BLOCK, VEC>(tile);
        if (p.vec_ok && tile_is_full<BLOCK, VEC>(tile, p.n)) {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
                __builtin_nontemporal_store(x[k], reinterpret_cast<f32x4*>(p.out + base + (uint64_t)k * 256u));
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint64_t i = base + (uint64_t)k * 256u + (uint64_t)c;
                    if (i < p.n) p.out[i] = x[k][c];
                }
        }
        if (!has_next) break;
#pragma unroll
        for (int k = 0; k < VEC; ++k) x[k] = y[k];
        tile = next;
        cur ^= 1u;
    }
}

#define GPS_FP_SHAPES(X) \
    X(256, 4)            \
    X(512, 8)            \
    X(1024, 4)           \
    X(1024, 8)

bool fp_shape_supported(TileShape s) {
#define X(B, V) \
    if (s.block == B && s.vec == V) return true;
    GPS_FP_SHAPES(X)
#undef X
    return false;
}

template <int B, int V, bool BF16, class Ops>
static hipError_t launch_fp_pp(const FpParams& p, hipStream_t stream) {
    const uint32_t resident = resident_workgroups<scan_fp_pp<B, V, BF16, Ops>, B>();
    if (resident == 0) return hipErrorInvalidDevice;
    const uint32_t grid = persistent_grid(p.num_tiles, resident);
    hipLaunchKernelGGL((scan_fp_pp<B, V, BF16, Ops>), dim3(grid), dim3(B), 0, stream, p);
    return hipGetLastError();
}

template <int B, int V, bool BF16, class Ops>
static hipError_t launch_fp_bl(const FpParams& p, hipStream_t stream) {
    const uint32_t resident = resident_workgroups<scan_fp_bl<B, V, BF16, Ops>, B>();
    if (resident == 0) return hipErrorInvalidDevice;
    const uint32_t grid = persistent_grid(p.num_tiles, resident);
    hipLaunchKernelGGL((scan_fp_bl<B, V, BF16, Ops>), dim3(grid), dim3(B), 0, stream, p);
    return hipGetLastError();
}

// Pair carries run the lane-blocked kernel; plain carries, with little
// arithmetic to save, the striped one.
template <int B, int V, bool BF16>
static hipError_t launch_fp_pp_carry(int carry, const FpParams& p, hipStream_t stream) {
    if (carry == kFpPair) return launch_fp_bl<B, V, BF16, PairCarry<true>>(p, stream);
    return launch_fp_pp<B, V, BF16, PairCarry<false>>(p, stream);
}

hipError_t launch_scan_fp(TileShape s, bool bf16, int carry, bool persistent, const void* in, float* out,
                          uint64_t n, uint32_t vec_ok, uint64_t* status, uint64_t* ticket, uint32_t* err,
                          hipStream_t stream) {
    const bool compensated = carry == kFpPair;
    FpParams p{};
    p.in = in;
    p.out = out;
    p.n = n;
    p.num_tiles = (uint32_t)((n + s.tile() - 1) / s.tile());
    p.vec_ok = vec_ok;
    p.status = status;
    p.ticket = ticket;
    p.err = err;
#define X(B, V)                                                                                                 \
    if (s.block == B && s.vec == V) {                                                                           \
        if (persistent)                                                                                         \
            return bf16 ? launch_fp_pp_carry<B, V, true>(carry, p, stream)                                      \
                        : launch_fp_pp_carry<B, V, false>(carry, p, stream);                                    \
        if (bf16 && compensated)                                                                                \
            hipLaunchKernelGGL((scan_fp<B, V, true, true>), dim3(p.num_tiles), dim3(B), 0, stream, p);       \
        else if (bf16)                                                                                          \
            hipLaunchKernelGGL((scan_fp<B, V, true, false>), dim3(p.num_tiles), dim3(B), 0, stream, p);      \
        else if (compensated)                                                                                   \
            hipLaunchKernelGGL((scan_fp<B, V, false, true>), dim3(p.num_tiles), dim3(B), 0, stream, p);      \
        else                                                                                                    \
            hipLaunchKernelGGL((scan_fp<B, V, false, false>), dim3(p.num_tiles), dim3(B), 0, stream, p);     \
        return hipGetLastError();                                                                               \
    }
    GPS_FP_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

template <int B, int V, bool BF16, bool COMP>
static hipError_t launch_segfp_pp(const SegFpParams& p, hipStream_t stream) {
    // Pair carries run the lane-blocked kernel; plain carries the striped one.
    if constexpr (COMP) {
        const uint32_t resident = resident_workgroups<segscan_fp_bl<B, V, BF16, COMP>, B>();
        if (resident == 0) return hipErrorInvalidDevice;
        const uint32_t grid = persistent_grid(p.num_tiles, resident);
        hipLaunchKernelGGL((segscan_fp_bl<B, V, BF16, COMP>), dim3(grid), dim3(B), 0, stream, p);
    } else {
        const uint32_t resident = resident_workgroups<segscan_fp_pp<B, V, BF16, COMP>, B>();
        if (resident == 0) return hipErrorInvalidDevice;
        const uint32_t grid = persistent_grid(p.num_tiles, resident);
        hipLaunchKernelGGL((segscan_fp_pp<B, V, BF16, COMP>), dim3(grid), dim3(B), 0, stream, p);
    }
    return hipGetLastError();
}

hipError_t launch_segscan_fp(TileShape s, bool bf16, bool compensated, bool persistent, const void* in, float* out,
                             const uint64_t* offsets, uint64_t num_segments, uint64_t n, uint32_t vec_ok,
                             uint64_t* status, uint64_t* ticket, uint32_t* err, hipStream_t stream) {
    SegFpParams p{};
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
#define X(B, V)                                                                                                  \
    if (s.block == B && s.vec == V) {                                                                            \
        if (persistent)                                                                                          \
            return bf16 ? (compensated ? launch_segfp_pp<B, V, true, true>(p, stream)                            \
                                       : launch_segfp_pp<B, V, true, false>(p, stream))                          \
                        : (compensated ? launch_segfp_pp<B, V, false, true>(p, stream)                           \
                                       : launch_segfp_pp<B, V, false, false>(p, stream));                        \
        if (bf16 && compensated)                                                                                 \
            hipLaunchKernelGGL((segscan_fp<B, V, true, true>), dim3(p.num_tiles), dim3(B), 0, stream, p);     \
        else if (bf16)                                                                                           \
            hipLaunchKernelGGL((segscan_fp<B, V, true, false>), dim3(p.num_tiles), dim3(B), 0, stream, p);    \
        else if (compensated)                                                                                    \
            hipLaunchKernelGGL((segscan_fp<B, V, false, true>), dim3(p.num_tiles), dim3(B), 0, stream, p);    \
        else                                                                                                     \
            hipLaunchKernelGGL((segscan_fp<B, V, false, false>), dim3(p.num_tiles), dim3(B), 0, stream, p);   \
        return hipGetLastError();                                                                                \
    }
    GPS_FP_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace gps
