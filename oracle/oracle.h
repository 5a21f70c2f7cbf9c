/*
 * CPU oracle for the uint32 / segmented / floating-point prefix sums.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: the test suite, the
 * smoke check and the benchmark's CPU-baseline leg call it to produce expected
 * outputs and reference timings. No product code path links or calls it; the
 * HIP library (gpuprefixsums_amd/libgps.so) never falls back to it.
 *
 * The reference ships no host-side scan. Its only element-wise scan oracle is
 * the single-thread validator loop in
 *   GPUPrefixSumsD3D12/Shaders/Utility.hlsl:99-125
 *   (copied in GPUPrefixSumsUnity/Shaders/Utility.compute:131-187):
 *       sum += in[i]; check out[i] == sum        (inclusive)
 *       check out[i] == sum; sum += in[i]        (exclusive)
 * in uint32 arithmetic, i.e. with wraparound mod 2^32. orc_scan_u32 and
 * orc_validate_u32 restate that loop. The closed forms used by the CUDA tests
 * (all-ones input -> i+1 / i, GPUPrefixSumsCUDA/UtilityKernels.cuh:22-38) are
 * restated by orc_count_ones_errors.
 *
 * Everything else here (segmented scan, FP64 scan, BF16 conversion, input
 * generators, the multi-threaded CPU baseline) has no reference counterpart;
 * each function states the published definition it follows and how the test
 * suite pins it.
 */
#ifndef GPS_ORACLE_H
#define GPS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- uint32 scan (Utility.hlsl:99-125) --------------------------------- */

/* Serial uint32 scan with wraparound. inclusive != 0: out[i] = in[0]+..+in[i];
 * else out[i] = in[0]+..+in[i-1] (out[0] = 0). out may alias in. */
void orc_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, int inclusive);

/* The reference's validator loop verbatim in meaning: counts i where
 * scan[i] differs from the serial running sum of in[] (Utility.hlsl:101-125). */
uint64_t orc_validate_u32(const uint32_t* in, const uint32_t* scan, uint64_t n, int inclusive);

/* Closed-form check of an all-ones scan (UtilityKernels.cuh:22-38):
 * counts i where scan[i] != i+1 (inclusive) or != i (exclusive), in uint32. */
uint64_t orc_count_ones_errors(const uint32_t* scan, uint64_t n, int inclusive);

/* Multi-threaded blocked two-pass scan (reduce per block, serial scan of the
 * block sums, rescan with carry). Same output as orc_scan_u32. Used only as
 * the all-cores CPU baseline. threads <= 0 means hardware concurrency. */
void orc_scan_u32_mt(const uint32_t* in, uint32_t* out, uint64_t n, int inclusive, int threads);

/* ---- segmented uint32 scan (no reference counterpart) ------------------ */

/* Segment s covers [offsets[s], offsets[s+1]); offsets[0] must be 0 and the
 * array non-decreasing; S = number of segments; n = offsets[S]. The running
 * sum restarts at every segment start. Returns 0, or -1 on bad offsets. */
int orc_segmented_scan_u32(const uint32_t* in, uint32_t* out, const uint64_t* offsets,
                           uint64_t num_segments, int inclusive);

/* ---- floating point (no reference counterpart) -------------------------- */

/* Inclusive scan of float32 input accumulated in float64 with Neumaier
 * compensation (the FP64 oracle). */
void orc_scan_f32_f64(const float* in, double* out, uint64_t n);
/* Same over bfloat16 input (raw uint16 bit patterns). */
void orc_scan_bf16_f64(const uint16_t* in, double* out, uint64_t n);
/* Running sum of |x| in float64 (error-bound denominators). */
void orc_abs_scan_f32_f64(const float* in, double* out, uint64_t n);
void orc_abs_scan_bf16_f64(const uint16_t* in, double* out, uint64_t n);

/* bfloat16 <-> float32 from the format definition (bf16 = top 16 bits of the
 * IEEE-754 binary32 pattern). f32 -> bf16 rounds to nearest, ties to even;
 * NaN stays NaN (quiet). */
float orc_bf16_to_f32(uint16_t b);
uint16_t orc_f32_to_bf16_rne(float f);

/* ---- input generators (fixed seeds, counter-based splitmix64) ----------- */

/* splitmix64 (Steele, Lea, Flood 2014; Vigna's reference C): element i of the
 * stream with seed s is mix(s + (i+1) * 0x9E3779B97F4A7C15). */
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i);

/* out[i] = top `bits` bits of orc_splitmix64_at(seed, first + i), 1 <= bits <= 32.
 * bits = 8 gives [0,255]; 10 gives [0,1023]; 32 the full uint32 range. */
void orc_gen_u32_bits(uint64_t seed, uint64_t first, uint32_t* out, uint64_t n, int bits);

/* Standard normal by Box-Muller in double from counters 2i and 2i+1:
 * u1 = ((r1 >> 11) + 1) * 2^-53, u2 = (r2 >> 11) * 2^-53,
 * z = sqrt(-2 ln u1) * cos(2 pi u2). */
double orc_normal_at(uint64_t seed, uint64_t i);
/* FP32 inputs ~ N(0,1): out[i] = (float)orc_normal_at(seed, i). */
void orc_gen_f32_normal(uint64_t seed, float* out, uint64_t n);
/* BF16 inputs with random sign and magnitude 10^U(-4,4): x = +-10^(8u-4) with
 * u = (r >> 11) * 2^-53 and the sign from bit 0 of r; rounded to FP32 (RNE)
 * then to BF16 (RNE). */
void orc_gen_bf16_logmag(uint64_t seed, uint16_t* out, uint64_t n);

/* Log-normal segment lengths: L_s = clip(round(exp(ln(median) + sigma * z_s)),
 * lo, hi) with z_s = orc_normal_at(seed, s). Writes offsets[0..S] (prefix of
 * lengths, offsets[0] = 0) and returns the total offsets[S]. */
uint64_t orc_gen_lognormal_offsets(uint64_t seed, uint64_t num_segments, double median,
                                   double sigma, uint64_t lo, uint64_t hi, uint64_t* offsets);

/* The D3D12 test generator, restated for reference-identical test inputs:
 * GPUPrefixSumsD3D12/Shaders/Utility.hlsl:37-66 (InitRandom). A hybrid
 * Tausworthe generator (GPU Gems 3, ch. 37: three Tausworthe steps and one
 * LCG step, XORed) seeded per thread from (4*tid + k) * seed, stepped once per
 * thread, masked to 10 bits. The shader runs 256x256 threads with stride
 * 65536, so element i gets the value of thread i mod 65536 (the shader
 * advances its state once per thread, not per element). */
void orc_gen_d3d12_taus(uint32_t seed, uint32_t* out, uint64_t n);

/* The Unity test generator: GPUPrefixSumsUnity/Shaders/Utility.compute:50-72
 * (InitRandom). The same hybrid Tausworthe steps, but the 65536 threads
 * (256 groups of 256) keep their state and step it once per uint4 element
 * they visit (uint4 element v belongs to thread v mod 65536), and the 32-bit
 * output is split into four bytes, low byte first, one per scan element.
 * Elements 4v..4v+3 of n (the last uint4 may be partial). The reference's
 * test harness never sets the seed (GPUPrefixSumsUnity/Runtime/Tests.cs
 * sets no "e_seed"), so its effective seed is 0: only the LCG component is
 * nonzero and every thread yields the same sequence. */
void orc_gen_unity_taus(uint32_t seed, uint32_t* out, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
