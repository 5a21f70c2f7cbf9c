/*
 * gps.h - C ABI of the MI355X (gfx950) prefix-sum library libgps.so.
 *
 * Plain pointers and sizes only: device pointers are HIP device addresses,
 * `stream` is a hipStream_t passed as void* (NULL = the null stream). Every
 * function returns a gps_status; gps_last_error() gives the message of the
 * calling thread's last failure. Nothing prints.
 *
 * Reference interfaces replaced (file:line in b0nes164/GPUPrefixSums):
 *   The reference's scan entry points are the private dispatcher methods
 *   DispatchKernelsInclusive / DispatchKernelsExclusive(uint32_t size)
 *     GPUPrefixSumsCUDA/ChainedScanDecoupledLookbackDispatcher.cuh:177-195
 *     GPUPrefixSumsCUDA/ReduceThenScanDispatcher.cuh:171-187
 *     GPUPrefixSumsCUDA/EmulatedDeadlockingDispatcher.cuh:159-166
 *   which scan the dispatcher-owned buffer in place after clearing the status
 *   words and the tile counter (ClearMemory, ...Dispatcher.cuh:172-175). Here
 *   the handle owns that scratch, the caller owns the data (in == out gives the
 *   reference's in-place behaviour), and no clear is needed per call.
 */
#ifndef GPS_H
#define GPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gps_status {
    GPS_OK = 0,
    GPS_ERR_INVALID = 1,     /* bad argument (null pointer, bad tuning, ...) */
    GPS_ERR_TOO_LARGE = 2,   /* n exceeds the handle's max_n */
    GPS_ERR_HIP = 3,         /* a HIP runtime call failed */
    GPS_ERR_UNSUPPORTED = 4, /* operation not offered by this handle's algorithm */
    GPS_ERR_DEVICE = 5       /* a device-side watchdog fired (see gps_device_error) */
} gps_status;

typedef enum gps_algo {
    GPS_ALGO_CSDL = 0,   /* chained scan with decoupled lookback
                            (ChainedScanDecoupledLookback.cuh:79-153) */
    GPS_ALGO_RTS = 1,    /* reduce-then-scan (ReduceThenScan.cuh:15-161) */
    GPS_ALGO_CSDLDF = 2  /* decoupled lookback with decoupled fallback, optional
                            emulated deadlock (EmulatedDeadlocking.cuh:26-168;
                            ...DecoupledFallback.hlsl:81-242) */
} gps_algo;

/* Runtime tuning; zero fields take defaults. Reference knobs: CSDL_WARPS /
 * UINT4_PER_THREAD (ChainedScanDecoupledLookbackDispatcher.cuh:13-14),
 * MAX_SPIN_COUNT / DEADLOCK_MASK (EmulatedDeadlocking.cuh:15-16). */
typedef struct gps_tuning {
    uint32_t block_threads;  /* 256, 512 or 1024 (whole wave64s) */
    uint32_t vec_per_thread; /* uint4 vectors per thread: 2, 4 or 8 (items = 4x) */
    uint32_t spin_limit;     /* CSDLDF: polls without progress before a fallback (default 4), for the
                                blocking tile's nearest successor; a waiter k tiles farther waits
                                (1 + min(k, 7)) x spin_limit polls; at most 65536 */
    uint32_t stall_period;   /* CSDLDF: stall tiles with index % period == 0 (0 = no emulation, >= 2);
                                with GPS_SCHEDULE_SPLIT, defer them to a second launch instead */
    uint32_t schedule;       /* CSDL, CSDLDF: GPS_SCHEDULE_* (0 = default) */
} gps_tuning;

/* CSDL workgroup schedules. TILE: one workgroup per tile, as the reference
 * launches (ChainedScanDecoupledLookbackDispatcher.cuh:192-194). PERSISTENT:
 * resident workgroups loop over ticketed tiles and load tile i+1 while tile
 * i's lookback resolves. CSDLDF handles take TILE (their default) or
 * PERSISTENT; RTS handles take neither. */
#define GPS_SCHEDULE_DEFAULT 0
#define GPS_SCHEDULE_TILE 1
#define GPS_SCHEDULE_PERSISTENT 2
/* CONTROL: persistent, with one wave per workgroup dedicated to tickets and
 * lookback; tiles are (block_threads - 64) * vec_per_thread * 4 elements. */
#define GPS_SCHEDULE_CONTROL 3
/* SPLIT (CSDLDF only, stall_period P >= 2): the deadlock emulation of the
 * reference's D3D12 path (GPUPrefixSumsD3D12/EmulatedDeadlock.cpp:97-118,
 * Shaders/EmulatedDeadlock.hlsl:183-253). Each scan is two launches on the
 * stream: the first runs every tile except those with index % P == 0, which
 * have not started when their successors look back, so every such tile with
 * a successor is reduced and inserted by a fallback; the second launch then
 * runs the deferred tiles. One workgroup per tile; no in-kernel stalls. */
#define GPS_SCHEDULE_SPLIT 4

typedef struct gps_handle gps_handle;

/* Library version string. */
const char* gps_version(void);
/* Message of this thread's last failed call ("" if none). */
const char* gps_last_error(void);

/* Creates a handle that owns the scratch for scans of up to max_n elements on
 * the current HIP device (ChainedScanDecoupledLookbackDispatcher.cuh:28-34).
 * tuning may be NULL. The ticket and status words (max_n / 8 + 256 bytes)
 * are in uncached device memory, which the library never returns to the
 * runtime: gps_destroy puts the block on a per-device free list that later
 * handles reuse. Setting GPS_STATUS_MEMORY=plain in the environment before
 * gps_create gives the handle ordinary device memory instead (freed by
 * gps_destroy; the scans run about 2% slower at 2^28 elements). */
int gps_create(int algo, uint64_t max_n, const gps_tuning* tuning, gps_handle** out);
/* Waits for the handle's device to finish its work, then releases the
 * handle (...Dispatcher.cuh:158-163). */
int gps_destroy(gps_handle* h);
/* Bytes of uncached status memory the library holds on all devices, and how
 * many of them are on the free lists (no reference counterpart). */
int gps_status_memory(uint64_t* held_bytes, uint64_t* free_bytes);
/* Effective tuning and tile size (elements per workgroup) of a handle. A
 * GPS_ALGO_CSDL handle created without a tile shape or schedule picks both
 * per call from n (one workgroup per tile, and smaller tiles, below 2^24
 * elements); for it this reports the configuration used from 2^24 up. */
int gps_get_tuning(const gps_handle* h, gps_tuning* out, uint32_t* tile_elems);

/* uint32 scans, wraparound mod 2^32, bit-exact with the serial loop of
 * GPUPrefixSumsD3D12/Shaders/Utility.hlsl:99-125. d_in may equal d_out.
 * Enqueued on `stream`; returns after the launch, not the completion.
 * Replaces DispatchKernelsInclusive (...Dispatcher.cuh:187-195) and
 * DispatchKernelsExclusive (...Dispatcher.cuh:177-185). */
int gps_inclusive_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream);
int gps_exclusive_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream);
/* Both in one entry point: inclusive != 0 selects the inclusive scan (the
 * form an FFI binds when it passes the direction as a flag). */
int gps_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* stream);

/* The same scan continued from a carry (no reference counterpart; it serves
 * chunked scans of arrays larger than one call, and the per-shard scan of a
 * multi-device scan): every output gets *d_carry_in added (mod 2^32), and
 * *d_carry_out receives the scan's total plus that carry. Both are device
 * words and either may be NULL (carry 0 / not written); they may be the same
 * word, so consecutive chunks chain through one word with no host sync.
 * Offered on GPS_ALGO_CSDL handles with the tile or persistent schedule
 * (a default handle's every choice); others return GPS_ERR_UNSUPPORTED. */
int gps_scan_u32_carry(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive,
                       const uint32_t* d_carry_in, uint32_t* d_carry_out, void* stream);

/* ---- batched segmented scan (no reference counterpart) ------------------- */

/* Scans num_segments segments stored back to back: segment s is
 * [d_offsets[s], d_offsets[s+1]) of d_in / d_out, and the running sum restarts
 * at every segment start (position 0 always starts one). d_offsets is a
 * device array of num_segments + 1 non-decreasing uint64 offsets with
 * d_offsets[num_segments] == n; empty segments are allowed. Exclusive output
 * puts 0 at each segment start. One kernel launch, no allocation and no
 * memset, so it may be captured into a graph. Needs a GPS_ALGO_CSDL handle. */
int gps_segmented_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, const uint64_t* d_offsets,
                           uint64_t num_segments, uint64_t n, int inclusive, void* stream);
/* The same result with one plain scan launch per non-empty segment, from
 * host-side offsets (the batched scan's baseline). */
int gps_segmented_scan_u32_per_segment(gps_handle* h, const uint32_t* d_in, uint32_t* d_out,
                                       const uint64_t* h_offsets, uint64_t num_segments, int inclusive,
                                       void* stream);

/* A hipGraph holding one gps_segmented_scan_u32 call with fixed arguments
 * (captured on a private stream). Replays reuse the handle's status words
 * through their epochs; after 2^30 - 1 replays call gps_reset first. */
typedef struct gps_graph gps_graph;
int gps_segmented_graph_create(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, const uint64_t* d_offsets,
                               uint64_t num_segments, uint64_t n, int inclusive, gps_graph** out);
int gps_graph_launch(gps_graph* g, void* stream);
int gps_graph_destroy(gps_graph* g);

/* ---- member-wise scan of 4-member uint32 structs ---------------------------
 * n structs of four uint32 members each (uint4 elements, 16-byte aligned
 * buffers, d_in == d_out allowed): for each member m independently,
 * out[i].m = sum of in[j].m over j <= i (inclusive) or j < i (exclusive),
 * mod 2^32. One launch; a status word per member per tile and a lookback that
 * takes a tile when all four agree. Replaces the reference's struct scan
 * (GPUPrefixSumsWebGPUapis/SharedShaders/csdldf_struct.wgsl:60-300, a vec4
 * scanned with per-member status words). Needs a GPS_ALGO_CSDL handle with
 * max_n >= n; a default handle picks shape and schedule from n as for the
 * plain scan. */
int gps_inclusive_scan_u32x4(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream);
int gps_exclusive_scan_u32x4(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream);

/* Occupancy estimate of the persistent scan, after the reference's
 * occupancy test variants (GPUPrefixSumsWebGPUapis/SharedShaders/TestVariants/
 * csdldf_occ.wgsl:70-82 and csdldf_struct_occ.wgsl, read back by
 * Dawn/main.cpp:663-672): an inclusive scan of d_in into d_out (n uint32, or
 * n structs when struct_scan != 0) launched with one workgroup per tile.
 * Workgroups that find the tickets already gone exit at once, so *measured
 * counts those that took a tile: the workgroups resident while the scan ran.
 * *resident is the grid the normal persistent launch uses (the HIP occupancy
 * query times the CU count). The scan's output is the ordinary result.
 * Synchronises `stream`. Needs a GPS_ALGO_CSDL handle. */
int gps_estimate_occupancy(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int struct_scan,
                           void* stream, uint32_t* measured, uint32_t* resident);

/* ---- survey of wave and block scan networks -------------------------------
 * The reference's survey (GPUPrefixSumsD3D12/Shaders/Survey.hlsl:49-855):
 * serial, wave and block Kogge-Stone, Sklansky, Brent-Kung-Blelloch and
 * raking scans, here as wave64 device functions run by a batch kernel over
 * independent segments of a fixed length per network. */
typedef enum gps_survey_alg {
    GPS_SURVEY_SERIAL = 0,                  /* one thread per 64-element segment */
    GPS_SURVEY_WAVE_KOGGE_STONE_LDS = 1,    /* 64: through LDS, no cross-lane ops */
    GPS_SURVEY_WAVE_KOGGE_STONE_SHUFFLE = 2,/* 64: ds_bpermute shuffles */
    GPS_SURVEY_WAVE_DPP = 3,                /* 64: DPP row_shr / row_bcast (the scans' own) */
    GPS_SURVEY_WAVE_RAKING = 4,             /* 1024: one wave, 16 DPP rounds with a carry */
    GPS_SURVEY_BLOCK_KOGGE_STONE = 5,       /* 256: LDS, one element per thread */
    GPS_SURVEY_BLOCK_SKLANSKY = 6,          /* 256 */
    GPS_SURVEY_BLOCK_BRENT_KUNG = 7,        /* 256: Blelloch upsweep / downsweep */
    GPS_SURVEY_BLOCK_WAVE_RAKING = 8,       /* 256: DPP wave scans + wave totals */
    GPS_SURVEY_BLOCK_WAVE_SKLANSKY = 9,     /* 256: DPP wave scans + Sklansky fans */
    GPS_SURVEY_TILE_REGISTERS = 10          /* 4096: the scans' register tile (256 x 4 uint4) */
} gps_survey_alg;
/* Segment length (elements) of a network. */
int gps_survey_segment_u32(int alg, uint32_t* segment);
/* Scans every segment of d_in into d_out independently (inclusive or
 * exclusive, mod 2^32); n must be a multiple of the segment length. 4-byte
 * aligned buffers (16-byte for GPS_SURVEY_TILE_REGISTERS). One launch. */
int gps_survey_scan_u32(int alg, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* stream);

/* ---- floating-point inclusive scans (no reference counterpart) ----------- */

/* FP32 or BF16 (raw uint16 bit patterns) input, FP32 output. `compensated`
 * selects how every partial sum above the element level (lane, wave, tile and
 * the inter-tile lookback chain) is carried:
 *   GPS_FP_PLAIN (0): one FP32 float, the plain chained scan;
 *   GPS_FP_PAIR  (1): a two-float pair combined with TwoSum, each output
 *                     rounded once.
 * Other values return GPS_ERR_INVALID. No FP64 arithmetic is used. One
 * kernel launch; needs a GPS_ALGO_CSDL handle. Runs the persistent kernel on
 * a handle with the persistent schedule (a default handle: above 2^24
 * elements), else one workgroup per tile. */
#define GPS_FP_PLAIN 0
#define GPS_FP_PAIR 1
int gps_inclusive_scan_f32(gps_handle* h, const float* d_in, float* d_out, uint64_t n, int compensated,
                           void* stream);
int gps_inclusive_scan_bf16_f32(gps_handle* h, const uint16_t* d_in, float* d_out, uint64_t n, int compensated,
                                void* stream);

/* ---- learned tuner: MLP cost model on the GPU (no reference counterpart) --- */

/* A 3-layer perceptron (in_dim -> hidden -> hidden -> 1, ReLU) predicting
 * log2 scan time from workload and configuration features, trained with MSE
 * and Adam by HIP kernels. Parameters are one flat FP32 array:
 * [W1 (hidden x in_dim), b1, W2 (hidden x hidden), b2, W3 (hidden), b3].
 * Inputs, targets and outputs are host arrays (the data set is small).
 * in_dim and hidden at most 4096; max_batch at most 65535 * 128 samples
 * (weight gradients sum the batch in chunks of 128 samples, added in chunk
 * order, so every run sums in the same order). */
typedef struct gps_mlp gps_mlp;
uint64_t gps_mlp_param_count(uint32_t in_dim, uint32_t hidden);
int gps_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t max_batch, gps_mlp** out);
int gps_mlp_destroy(gps_mlp* m);
/* Loads parameters and resets the optimiser state. */
int gps_mlp_set_params(gps_mlp* m, const float* h_params);
int gps_mlp_get_params(gps_mlp* m, float* h_params);
/* y[batch] = model(x[batch x in_dim]). */
int gps_mlp_forward(gps_mlp* m, const float* h_x, uint32_t batch, float* h_y);
/* Mean squared error over the batch and its gradient w.r.t. every parameter
 * (h_grad: param_count floats; either output may be NULL). */
int gps_mlp_loss_grad(gps_mlp* m, const float* h_x, const float* h_t, uint32_t batch, float* h_loss,
                      float* h_grad);
/* Full-batch AdamW (beta1 0.9, beta2 0.999, eps 1e-8) for `epochs` steps;
 * h_loss_curve (epochs floats, may be NULL) receives the loss before each step. */
int gps_mlp_train(gps_mlp* m, const float* h_x, const float* h_t, uint32_t batch, uint32_t epochs, float lr,
                  float weight_decay, float* h_loss_curve);

/* Batched segmented FP inclusive scan in one launch over the concatenated
 * segments: segment s covers [d_offsets[s], d_offsets[s+1]) (device array of
 * num_segments + 1 entries, d_offsets[0] = 0, d_offsets[S] = n; empty
 * segments allowed). Same input/output types as the plain FP scans;
 * compensated is GPS_FP_PLAIN or GPS_FP_PAIR. The carry restarts at every segment head at every
 * level (lane, wave, tile, lookback). Needs a GPS_ALGO_CSDL handle. Runs the
 * persistent kernel on a handle with the persistent schedule (a default
 * handle: above 2^22 elements, at 1024 x 8), else one workgroup per tile. */
int gps_segmented_scan_f32(gps_handle* h, const float* d_in, float* d_out, const uint64_t* d_offsets,
                           uint64_t num_segments, uint64_t n, int compensated, void* stream);
int gps_segmented_scan_bf16_f32(gps_handle* h, const uint16_t* d_in, float* d_out, const uint64_t* d_offsets,
                                uint64_t num_segments, uint64_t n, int compensated, void* stream);

/* One FP scan launch per non-empty segment (host offsets; d_in is float or
 * bf16 per `bf16`): the per-segment baseline of the batched FP form. */
int gps_inclusive_scan_fp_per_segment(gps_handle* h, const void* d_in, float* d_out, const uint64_t* h_offsets,
                                      uint64_t num_segments, int bf16, int compensated, void* stream);

/* CSDLDF fallback counters accumulated since the last reset:
 * [0] status polls, [1] fallbacks (redundant tile reductions), [2] successful
 * CAS insertions, [3] emulated stalls that hit their watchdog.
 * Synchronises the handle's device. Cf. TestVariants/csdldf_stats.wgsl:285-290. */
int gps_get_counters(gps_handle* h, uint64_t out[4]);
int gps_reset_counters(gps_handle* h);
/* Device error bits raised by watchdogs since the last call (and clears
 * them): 1 = a lookback exceeded its poll budget, 2 = an emulated stall did.
 * Synchronises the device. When any bit was set it also re-arms the handle's
 * status words and ticket word (as gps_reset does), because the failed launch
 * may have left tiles unpublished or tickets untaken; the next launch then
 * starts clean. */
int gps_device_error(gps_handle* h, uint32_t* bits);
/* Re-arms the status words (epoch wrap after 2^30-1 launches; the direct entry
 * points do this themselves, graph replays must call it). */
int gps_reset(gps_handle* h, void* stream);

/* Allocation-free uint32 scan (no counterpart in the reference). For callers
 * that cannot allocate device memory: code under stream capture, serving
 * loops, harnesses that forbid hipMalloc. Same result as gps_scan_u32 on a
 * default GPS_ALGO_CSDL handle, but with no handle: the scratch (status slots,
 * ticket and error words) is a static device array of the library, 1 MiB +
 * 256 B per GPU, held until the process ends; the call allocates nothing and
 * launches one kernel (two on the first call on a device, which also re-arms
 * the scratch). n at most 2^28 (GPS_ERR_TOO_LARGE beyond). Every call on a
 * GPU shares that GPU's scratch: calls that overlap in time on one GPU (other
 * streams, other host threads) race and must be serialised by the caller.
 * Device watchdog bits are not reported (no handle to read them from). */
int gps_scan_u32_noalloc(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive, void* stream);

/* Times `iters` back-to-back calls of the handle's scan on `stream` with a
 * hipEvent pair around each call; writes per-call milliseconds to ms[iters].
 * Cf. BatchTimingInclusive (...Dispatcher.cuh:116-156). */
int gps_time_scan_u32(gps_handle* h, const uint32_t* d_in, uint32_t* d_out, uint64_t n, int inclusive,
                      void* stream, uint32_t iters, float* ms);

/* ---- utility kernels (UtilityKernels.cuh:15-38 and a random fill) -------- */

int gps_fill_u32(uint32_t* d_ptr, uint64_t n, uint32_t value, void* stream);
/* Streaming copy with the scans' tile geometry (read n + write n): the
 * practical bandwidth ceiling the scans are measured against. No reference
 * counterpart. */
int gps_copy_u32(uint32_t* d_dst, const uint32_t* d_src, uint64_t n, void* stream);
/* d_ptr[i] = top `bits` bits of splitmix64(seed, first + i) (the same stream
 * the CPU oracle generates). */
int gps_fill_random_u32(uint32_t* d_ptr, uint64_t n, uint64_t seed, uint64_t first, uint32_t bits,
                        void* stream);
/* Counts mismatches on the device and returns them (synchronises `stream`):
 * all-ones closed form (UtilityKernels.cuh:22-38) and the parallel form of the
 * serial validator (Utility.hlsl:99-125). */
int gps_validate_ones_u32(const uint32_t* d_scan, uint64_t n, int inclusive, void* stream, uint64_t* errors);
int gps_validate_scan_u32(const uint32_t* d_in, const uint32_t* d_scan, uint64_t n, int inclusive,
                          void* stream, uint64_t* errors);

/* ---- device unit kernel for the wave64 primitives -------------------------
 * The reference's warp primitives (Utils.cuh:22-71: InclusiveWarpScan,
 * WarpReduceSum) are wave64 DPP routines here. This runs them on every run of
 * 64 consecutive elements (a tail run zero-padded) so tests can check them in
 * isolation: inclusive scan, run total in every lane, or exclusive scan. */
enum { GPS_WAVE_INCLUSIVE_SCAN = 0, GPS_WAVE_SUM = 1, GPS_WAVE_EXCLUSIVE_SCAN = 2 };
int gps_wave_primitive_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, int op, void* stream);

/* ---- reference dispatcher self-tests (C++ classes in gps_dispatchers.hpp) - */

/* Runs <Dispatcher>::TestAllInclusive / TestAllExclusive for algo with a
 * dispatcher of max_size elements; writes tests passed and tests run.
 * (ChainedScanDecoupledLookbackDispatcher.cuh:38-114) */
int gps_dispatcher_test_all(int algo, uint32_t max_size, int inclusive, uint32_t* passed, uint32_t* total);

#ifdef __cplusplus
}
#endif
#endif /* GPS_H */
