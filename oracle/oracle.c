/*
 * CPU oracle (test infrastructure only; see oracle.h). Plain C11 + pthreads.
 * Built by oracle/Makefile into oracle/liboracle.so.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ---- uint32 scan -------------------------------------------------------- */

/* Utility.hlsl:101-111 (inclusive: add, then compare) and :115-125
 * (exclusive: compare, then add), restated as a producer of the expected
 * output instead of a comparer. uint32_t arithmetic wraps mod 2^32 exactly as
 * the HLSL `uint sum` does. */
void orc_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, int inclusive) {
    uint32_t sum = 0;
    if (inclusive) {
        for (uint64_t i = 0; i < n; ++i) {
            sum += in[i];
            out[i] = sum;
        }
    } else {
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t x = in[i];
            out[i] = sum;
            sum += x;
        }
    }
}

/* Utility.hlsl:99-125: one running sum, one error increment per mismatch. */
uint64_t orc_validate_u32(const uint32_t* in, const uint32_t* scan, uint64_t n, int inclusive) {
    uint32_t sum = 0;
    uint64_t errors = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (inclusive) {
            sum += in[i];
            if (scan[i] != sum) ++errors;
        } else {
            if (scan[i] != sum) ++errors;
            sum += in[i];
        }
    }
    return errors;
}

/* UtilityKernels.cuh:22-38: ValidateInclusive checks scan[i] != i + 1,
 * ValidateExclusive checks scan[i] != i (uint32 comparisons). */
uint64_t orc_count_ones_errors(const uint32_t* scan, uint64_t n, int inclusive) {
    uint64_t errors = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t expect = (uint32_t)i + (inclusive ? 1u : 0u);
        if (scan[i] != expect) ++errors;
    }
    return errors;
}

typedef struct {
    const uint32_t* in;
    uint32_t* out;
    uint64_t begin, end;
    uint32_t carry; /* pass 2 input */
    uint32_t total; /* pass 1 output */
    int inclusive;
    int pass;
} mt_job;

static void* mt_worker(void* arg) {
    mt_job* j = (mt_job*)arg;
    if (j->pass == 1) {
        uint32_t s = 0;
        for (uint64_t i = j->begin; i < j->end; ++i) s += j->in[i];
        j->total = s;
    } else {
        uint32_t sum = j->carry;
        if (j->inclusive) {
            for (uint64_t i = j->begin; i < j->end; ++i) {
                sum += j->in[i];
                j->out[i] = sum;
            }
        } else {
            for (uint64_t i = j->begin; i < j->end; ++i) {
                const uint32_t x = j->in[i];
                j->out[i] = sum;
                sum += x;
            }
        }
    }
    return NULL;
}

void orc_scan_u32_mt(const uint32_t* in, uint32_t* out, uint64_t n, int inclusive, int threads) {
    if (threads <= 0) {
        long hc = sysconf(_SC_NPROCESSORS_ONLN);
        threads = hc > 0 ? (int)hc : 1;
    }
    if (threads > 256) threads = 256;
    if ((uint64_t)threads > n) threads = n ? (int)n : 1;
    mt_job jobs[256];
    pthread_t tids[256];
    const uint64_t chunk = (n + (uint64_t)threads - 1) / (uint64_t)threads;
    for (int t = 0; t < threads; ++t) {
        uint64_t b = chunk * (uint64_t)t, e = b + chunk;
        if (b > n) b = n;
        if (e > n) e = n;
        jobs[t] = (mt_job){in, out, b, e, 0, 0, inclusive, 1};
        pthread_create(&tids[t], NULL, mt_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    uint32_t carry = 0;
    for (int t = 0; t < threads; ++t) {
        jobs[t].carry = carry;
        carry += jobs[t].total;
        jobs[t].pass = 2;
    }
    for (int t = 0; t < threads; ++t) pthread_create(&tids[t], NULL, mt_worker, &jobs[t]);
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
}

/* ---- segmented ---------------------------------------------------------- */

int orc_segmented_scan_u32(const uint32_t* in, uint32_t* out, const uint64_t* offsets,
                           uint64_t num_segments, int inclusive) {
    if (offsets[0] != 0) return -1;
    for (uint64_t s = 0; s < num_segments; ++s) {
        if (offsets[s + 1] < offsets[s]) return -1;
        uint32_t sum = 0;
        for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i) {
            const uint32_t x = in[i];
            if (inclusive) {
                sum += x;
                out[i] = sum;
            } else {
                out[i] = sum;
                sum += x;
            }
        }
    }
    return 0;
}

/* ---- floating point ------------------------------------------------------ */

/* Neumaier's improved Kahan-Babuska summation (Neumaier 1974), in double. */
#define NEUMAIER_STEP(x)                          \
    do {                                          \
        const double v_ = (x);                    \
        const double t_ = s + v_;                 \
        if (fabs(s) >= fabs(v_))                  \
            c += (s - t_) + v_;                   \
        else                                      \
            c += (v_ - t_) + s;                   \
        s = t_;                                   \
    } while (0)

void orc_scan_f32_f64(const float* in, double* out, uint64_t n) {
    double s = 0.0, c = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        NEUMAIER_STEP((double)in[i]);
        out[i] = s + c;
    }
}

void orc_scan_bf16_f64(const uint16_t* in, double* out, uint64_t n) {
    double s = 0.0, c = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        NEUMAIER_STEP((double)orc_bf16_to_f32(in[i]));
        out[i] = s + c;
    }
}

void orc_abs_scan_f32_f64(const float* in, double* out, uint64_t n) {
    double s = 0.0, c = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        NEUMAIER_STEP(fabs((double)in[i]));
        out[i] = s + c;
    }
}

void orc_abs_scan_bf16_f64(const uint16_t* in, double* out, uint64_t n) {
    double s = 0.0, c = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        NEUMAIER_STEP(fabs((double)orc_bf16_to_f32(in[i])));
        out[i] = s + c;
    }
}

/* bfloat16 is the upper half of a binary32 pattern. */
float orc_bf16_to_f32(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

/* Round-to-nearest-even onto the upper 16 bits: add 0x7FFF plus the lowest
 * kept bit, then truncate. Carry out of the mantissa correctly bumps the
 * exponent (and overflows to infinity at the top). NaNs are kept quiet. */
uint16_t orc_f32_to_bf16_rne(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x0040u);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return (uint16_t)(u >> 16);
}

/* ---- generators ---------------------------------------------------------- */

uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_gen_u32_bits(uint64_t seed, uint64_t first, uint32_t* out, uint64_t n, int bits) {
    if (bits < 1) bits = 1;
    if (bits > 32) bits = 32;
    const int shift = 64 - bits;
    for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)(orc_splitmix64_at(seed, first + i) >> shift);
}

double orc_normal_at(uint64_t seed, uint64_t i) {
    const uint64_t r1 = orc_splitmix64_at(seed, 2 * i);
    const uint64_t r2 = orc_splitmix64_at(seed, 2 * i + 1);
    const double u1 = (double)((r1 >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(r2 >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

void orc_gen_f32_normal(uint64_t seed, float* out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) out[i] = (float)orc_normal_at(seed, i);
}

void orc_gen_bf16_logmag(uint64_t seed, uint16_t* out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t r = orc_splitmix64_at(seed, i);
        const double u = (double)(r >> 11) * 0x1.0p-53;
        double x = pow(10.0, 8.0 * u - 4.0);
        if (r & 1u) x = -x;
        out[i] = orc_f32_to_bf16_rne((float)x);
    }
}

uint64_t orc_gen_lognormal_offsets(uint64_t seed, uint64_t num_segments, double median,
                                   double sigma, uint64_t lo, uint64_t hi, uint64_t* offsets) {
    const double mu = log(median);
    offsets[0] = 0;
    for (uint64_t s = 0; s < num_segments; ++s) {
        double len = round(exp(mu + sigma * orc_normal_at(seed, s)));
        if (len < (double)lo) len = (double)lo;
        if (len > (double)hi) len = (double)hi;
        offsets[s + 1] = offsets[s] + (uint64_t)len;
    }
    return offsets[num_segments];
}

void orc_gen_d3d12_taus(uint32_t seed, uint32_t* out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t tid = (uint32_t)(i & 0xFFFFu);
        uint32_t a = (tid << 2) * seed;
        uint32_t b = ((tid << 2) + 1u) * seed;
        uint32_t c = ((tid << 2) + 2u) * seed;
        uint32_t d = ((tid << 2) + 3u) * seed;
        a = ((a & 0xFFFFFFFEu) << 12) ^ (((a << 13) ^ a) >> 19);
        b = ((b & 0xFFFFFFF8u) << 4) ^ (((b << 2) ^ b) >> 25);
        c = ((c & 0xFFFFFFF0u) << 17) ^ (((c << 3) ^ c) >> 11);
        d = d * 1664525u + 1013904223u;
        out[i] = (a ^ b ^ c ^ d) & 1023u;
    }
}

void orc_gen_unity_taus(uint32_t seed, uint32_t* out, uint64_t n) {
    const uint64_t nvec = (n + 3u) / 4u;  /* uint4 elements, VectorizedSize(n) */
    for (uint32_t tid = 0; tid < 65536u && tid < nvec; ++tid) {
        uint32_t a = (tid << 2) * seed;
        uint32_t b = ((tid << 2) + 1u) * seed;
        uint32_t c = ((tid << 2) + 2u) * seed;
        uint32_t d = ((tid << 2) + 3u) * seed;
        for (uint64_t v = tid; v < nvec; v += 65536u) {
            a = ((a & 0xFFFFFFFEu) << 12) ^ (((a << 13) ^ a) >> 19);
            b = ((b & 0xFFFFFFF8u) << 4) ^ (((b << 2) ^ b) >> 25);
            c = ((c & 0xFFFFFFF0u) << 17) ^ (((c << 3) ^ c) >> 11);
            d = d * 1664525u + 1013904223u;
            const uint32_t t = a ^ b ^ c ^ d;
            for (uint32_t k = 0; k < 4u; ++k) {
                const uint64_t i = 4u * v + k;
                if (i < n) out[i] = (t >> (8u * k)) & 255u;
            }
        }
    }
}
