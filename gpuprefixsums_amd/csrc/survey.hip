// Survey of wave- and block-level scan networks on CDNA4 (wave64), as device
// building blocks with a batch kernel that scans independent fixed-length
// segments, so each network can be tested against the oracle and timed in
// place.
//
// Reference counterpart: GPUPrefixSumsD3D12/Shaders/Survey.hlsl:49-855 (serial,
// wave and block Kogge-Stone, Sklansky, Brent-Kung-Blelloch, raking
// reduce-scan, with and without wave intrinsics). Those scan one buffer in
// place with one workgroup; here every network is a device function over
// registers or LDS, the batch kernel runs it over many segments at once, and
// the wave-intrinsic forms use gfx950 DPP (row_shr, row_bcast) or ds_bpermute.
//
// Segment geometry (elements), fixed per network:
//   serial                  64 (one thread per segment)
//   wave networks           64 (one wave per segment, four per workgroup)
//   wave raking            1024 (one wave per segment: 16 rounds of 64)
//   block networks          256 (one 256-thread workgroup per segment)
//   register tile          4096 (256 threads x 4 uint4: the scans' own tile)
#include "gps_internal.hpp"
#include "tile_u32.hpp"

namespace gps {

namespace {

constexpr uint32_t kBlock = 256;

__host__ __device__ constexpr uint32_t segment_of(int alg) {
    return alg == kSurveySerial ? 64u
           : alg <= kSurveyWaveDpp ? 64u
           : alg == kSurveyWaveRaking ? 1024u
           : alg <= kSurveyBlockWaveSklansky ? 256u
                                             : 4096u;
}
// Elements one workgroup covers.
__host__ __device__ constexpr uint32_t workgroup_elems(int alg) {
    return alg == kSurveySerial ? kBlock * 64u
           : alg <= kSurveyWaveDpp ? kBlock
           : alg == kSurveyWaveRaking ? 4u * 1024u
           : alg <= kSurveyBlockWaveSklansky ? kBlock
                                             : 4096u;
}

// Branch-free element load: lanes past n read element 0 and take 0, so no
// divergent region precedes a DPP step or a barrier.
__device__ __forceinline__ uint32_t load_or_zero(const uint32_t* in, uint64_t i, uint64_t n) {
    const uint32_t v = in[i < n ? i : 0];
    return i < n ? v : 0u;
}

// ---- wave networks (64 lanes, one element each) ------------------------------

// Kogge-Stone through LDS, no cross-lane instructions: log2(64) rounds of
// read-neighbour / add, with the workgroup barrier between a round's reads
// and writes (Survey.hlsl:71-118).
__device__ __forceinline__ uint32_t wave_ks_lds(uint32_t x, uint32_t* s_buf) {
    const uint32_t tid = threadIdx.x, lane = lane_id();
    s_buf[tid] = x;
    __syncthreads();
#pragma unroll
    for (uint32_t d = 1; d < kWave; d <<= 1) {
        const uint32_t t = lane >= d ? s_buf[tid - d] : 0u;
        __syncthreads();
        s_buf[tid] += t;
        __syncthreads();
    }
    return s_buf[tid];
}

// Kogge-Stone with ds_bpermute shuffles (Survey.hlsl:120-165).
__device__ __forceinline__ uint32_t wave_ks_shuffle(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t d = 1; d < kWave; d <<= 1) {
        const uint32_t src = lane >= d ? lane - d : lane;
        const uint32_t t = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)x);
        x += lane >= d ? t : 0u;
    }
    return x;
}

// ---- block networks (256 threads, one element each, in LDS) ------------------

// Kogge-Stone over the workgroup (Survey.hlsl:229-267).
__device__ __forceinline__ uint32_t block_ks(uint32_t x, uint32_t* s_buf) {
    const uint32_t tid = threadIdx.x;
    s_buf[tid] = x;
    __syncthreads();
#pragma unroll
    for (uint32_t d = 1; d < kBlock; d <<= 1) {
        const uint32_t t = tid >= d ? s_buf[tid - d] : 0u;
        __syncthreads();
        s_buf[tid] += t;
        __syncthreads();
    }
    return s_buf[tid];
}

// Sklansky: at level k every element with bit k set adds the last element of
// the preceding 2^k-run; readers and writers of a level are disjoint
// (Survey.hlsl:269-300).
__device__ __forceinline__ uint32_t block_sklansky(uint32_t x, uint32_t* s_buf) {
    const uint32_t tid = threadIdx.x;
    s_buf[tid] = x;
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; (1u << k) < kBlock; ++k) {
        if (tid & (1u << k)) s_buf[tid] += s_buf[((tid >> k) << k) - 1u];
        __syncthreads();
    }
    return s_buf[tid];
}

// Brent-Kung / Blelloch, work-efficient: upsweep to the total, clear the
// root, downsweep swapping and adding; gives the exclusive scan
// (Survey.hlsl:302-362).
__device__ __forceinline__ uint32_t block_brent_kung_excl(uint32_t x, uint32_t* s_buf) {
    const uint32_t tid = threadIdx.x;
    s_buf[tid] = x;
    __syncthreads();
#pragma unroll
    for (uint32_t d = 1; d < kBlock; d <<= 1) {
        if (tid < kBlock / (2u * d)) {
            const uint32_t ai = d * (2u * tid + 1u) - 1u, bi = d * (2u * tid + 2u) - 1u;
            s_buf[bi] += s_buf[ai];
        }
        __syncthreads();
    }
    if (tid == 0) s_buf[kBlock - 1] = 0u;
    __syncthreads();
#pragma unroll
    for (uint32_t d = kBlock / 2; d >= 1; d >>= 1) {
        if (tid < kBlock / (2u * d)) {
            const uint32_t ai = d * (2u * tid + 1u) - 1u, bi = d * (2u * tid + 2u) - 1u;
            const uint32_t t = s_buf[ai];
            s_buf[ai] = s_buf[bi];
            s_buf[bi] += t;
        }
        __syncthreads();
    }
    return s_buf[tid];
}

// Raking over waves: DPP wave scans, the four wave totals scanned by one DPP
// pass in every wave (Survey.hlsl:679-714).
__device__ __forceinline__ uint32_t block_wave_raking(uint32_t x, uint32_t* s_wave) {
    const uint32_t wave = wave_index();
    const uint32_t s = wave_incl_scan(x);
    if (lane_id() == kWave - 1) s_wave[wave] = s;
    __syncthreads();
    uint32_t wave_prefix, total;
    combine_waves<kBlock / kWave>(s_wave, wave, wave_prefix, total);
    return s + wave_prefix;
}

// Sklansky fans across wave boundaries after DPP wave scans
// (Survey.hlsl:592-631).
__device__ __forceinline__ uint32_t block_wave_sklansky(uint32_t x, uint32_t* s_buf) {
    const uint32_t tid = threadIdx.x;
    s_buf[tid] = wave_incl_scan(x);
    __syncthreads();
#pragma unroll
    for (uint32_t k = 6; (1u << k) < kBlock; ++k) {
        if (tid & (1u << k)) s_buf[tid] += s_buf[((tid >> k) << k) - 1u];
        __syncthreads();
    }
    return s_buf[tid];
}

template <int ALG, bool INCL>
__global__ __launch_bounds__(kBlock) void survey_kernel(const uint32_t* in, uint32_t* out, uint64_t n) {
    __shared__ uint32_t s_buf[kBlock];
    const uint32_t tid = threadIdx.x, lane = lane_id();
    const uint64_t wg_base = (uint64_t)blockIdx.x * workgroup_elems(ALG);

    if constexpr (ALG == kSurveySerial) {
        const uint64_t base = wg_base + (uint64_t)tid * 64u;
        if (base < n) {
            uint32_t s = 0;
            for (uint32_t j = 0; j < 64u; ++j) {
                const uint32_t x = in[base + j];
                s += x;
                out[base + j] = INCL ? s : s - x;
            }
        }
    } else if constexpr (ALG == kSurveyWaveRaking) {
        const uint64_t base = wg_base + (uint64_t)wave_index() * 1024u;
        uint32_t carry = 0;
        for (uint32_t r = 0; r < 16u; ++r) {
            const uint64_t i = base + r * 64u + lane;
            const uint32_t x = load_or_zero(in, i, n);
            const uint32_t s = wave_incl_scan(x) + carry;
            if (i < n) out[i] = INCL ? s : s - x;
            carry = readlane(s, 63);
        }
    } else if constexpr (ALG == kSurveyTileRegisters) {
        using C = TileCfg<kBlock, 4>;
        __shared__ uint32_t s_wave[C::kWaves];
        u32x4 d[4];
        const uint64_t base = thread_base<kBlock, 4>(blockIdx.x);
        if (wg_base + C::kTile <= n)
            load_vec<4>(d, in + base);
        else
            load_guarded<4>(d, in, base, n);
        const uint32_t run = wave_scan_run<4, INCL>(d);
        if (lane == 0) s_wave[wave_index()] = run;
        __syncthreads();
        uint32_t wave_prefix, total;
        combine_waves<C::kWaves>(s_wave, wave_index(), wave_prefix, total);
        if (wg_base + C::kTile <= n)
            store_vec<4>(out + base, d, wave_prefix);
        else
            store_guarded<4>(out, d, wave_prefix, base, n);
    } else {
        const uint64_t i = wg_base + tid;
        const uint32_t x = load_or_zero(in, i, n);
        uint32_t r;
        if constexpr (ALG == kSurveyWaveKoggeStoneLds) {
            r = wave_ks_lds(x, s_buf);
        } else if constexpr (ALG == kSurveyWaveKoggeStoneShuffle) {
            r = wave_ks_shuffle(x);
        } else if constexpr (ALG == kSurveyWaveDpp) {
            r = wave_incl_scan(x);
        } else if constexpr (ALG == kSurveyBlockKoggeStone) {
            r = block_ks(x, s_buf);
        } else if constexpr (ALG == kSurveyBlockSklansky) {
            r = block_sklansky(x, s_buf);
        } else if constexpr (ALG == kSurveyBlockBrentKung) {
            r = block_brent_kung_excl(x, s_buf) + x;  // inclusive from exclusive
        } else if constexpr (ALG == kSurveyBlockWaveRaking) {
            r = block_wave_raking(x, s_buf);
        } else {
            r = block_wave_sklansky(x, s_buf);
        }
        if (i < n) out[i] = INCL ? r : r - x;
    }
}

}  // namespace

uint32_t survey_segment(int alg) { return alg >= 0 && alg < kSurveyCount ? segment_of(alg) : 0u; }

hipError_t launch_survey_u32(int alg, bool inclusive, const uint32_t* in, uint32_t* out, uint64_t n,
                             hipStream_t stream) {
    if (alg < 0 || alg >= kSurveyCount) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + workgroup_elems(alg) - 1) / workgroup_elems(alg);
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
#define GPS_SURVEY_CASE(A)                                                                                    \
    case A:                                                                                                   \
        if (inclusive)                                                                                        \
            hipLaunchKernelGGL((survey_kernel<A, true>), dim3((uint32_t)grid), dim3(kBlock), 0, stream, in, out, n); \
        else                                                                                                  \
            hipLaunchKernelGGL((survey_kernel<A, false>), dim3((uint32_t)grid), dim3(kBlock), 0, stream, in, out, n); \
        break;
    switch (alg) {
        GPS_SURVEY_CASE(kSurveySerial)
        GPS_SURVEY_CASE(kSurveyWaveKoggeStoneLds)
        GPS_SURVEY_CASE(kSurveyWaveKoggeStoneShuffle)
        GPS_SURVEY_CASE(kSurveyWaveDpp)
        GPS_SURVEY_CASE(kSurveyWaveRaking)
        GPS_SURVEY_CASE(kSurveyBlockKoggeStone)
        GPS_SURVEY_CASE(kSurveyBlockSklansky)
        GPS_SURVEY_CASE(kSurveyBlockBrentKung)
        GPS_SURVEY_CASE(kSurveyBlockWaveRaking)
        GPS_SURVEY_CASE(kSurveyBlockWaveSklansky)
        GPS_SURVEY_CASE(kSurveyTileRegisters)
        default:
            return hipErrorInvalidValue;
    }
#undef GPS_SURVEY_CASE
    return hipGetLastError();
}

}  // namespace gps
