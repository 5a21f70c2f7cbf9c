/* gps_ceiling.h - memory-ceiling probes for the scan's roofline, in their own
 * library (libgps_ceiling.so, which links libgps.so for its error reporting).
 * Benchmark only: the product library libgps.so carries none of them, and
 * they have no counterpart in the reference.
 *
 * Each probe moves `bytes` through HBM with one access pattern: one-shot
 * copies at several tile shapes, ticketed persistent copies at one workgroup
 * per CU (the persistent scan's schedule without the scan), a copy staged
 * through LDS with direct-to-LDS loads, one-direction streams, and the
 * runtime's own copy and fill. The single-pass scan reads n words and writes
 * n words, so a copy of the same bytes bounds it from above; these probes
 * show which access patterns move those bytes fastest on this part.
 *
 * Return codes are gps_status values; messages are read with gps_last_error()
 * from libgps.so. */
#ifndef GPS_CEILING_H
#define GPS_CEILING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gps_ceiling_kind {
    GPS_CEIL_COPY_64X1 = 0,       /* one-shot copy, 64 threads x 1 uint4 (1 KiB per workgroup) */
    GPS_CEIL_COPY_256X4 = 1,      /* one-shot copy, 256 threads x 4 uint4 (16 KiB) */
    GPS_CEIL_COPY_1024X8 = 2,     /* one-shot copy at the scan's tile (128 KiB) */
    GPS_CEIL_TICKET_1024X4 = 3,   /* ticketed persistent copy, 64 KiB tiles, one workgroup per CU */
    GPS_CEIL_TICKET_1024X8 = 4,   /* the same at the scan's 128 KiB tile */
    GPS_CEIL_COPY_LDS_256X4 = 5,  /* one-shot copy staged through LDS by direct-to-LDS loads */
    GPS_CEIL_READ_1024X1 = 6,     /* read-only stream of `bytes` from src */
    GPS_CEIL_WRITE_256X1 = 7,     /* write-only stream of `bytes` to dst */
    GPS_CEIL_MEMCPY = 8,          /* hipMemcpyAsync device to device */
    GPS_CEIL_MEMSET = 9,          /* hipMemsetD32Async */
    GPS_CEIL_TICKET_1024X4_STATUS = 10, /* ticketed copy 1024x4 plus the scan's status traffic per tile:
                                           an aggregate and an inclusive word published and one 64-word
                                           lookback poll, all relaxed device-scope atomics, no waits */
    GPS_CEIL_TICKET_1024X8_STATUS = 11, /* the same at 1024x8 */
    GPS_CEIL_TICKET_1024X4_STATUS_UC = 12, /* kind 10 with its ticket and status words in uncached
                                              device memory held by the library, as the scans keep
                                              theirs (no caller scratch) */
    GPS_CEIL_TICKET_1024X8_STATUS_UC = 13, /* the same at 1024x8 */
    GPS_CEIL_KINDS = 14
};

/* Bytes per tile of `kind`: `bytes` passed to gps_ceiling_run must be a
 * multiple of it. */
int gps_ceiling_granule(int kind, uint64_t* granule);

/* Short name of `kind` (static storage). */
const char* gps_ceiling_name(int kind);

/* Device scratch bytes `kind` needs for `bytes` of data: 0 for the one-shot
 * kinds and the uncached-status kinds; for the other ticketed kinds a ticket
 * word (256 bytes) plus, for the status kinds, one 128-byte status slot per
 * tile. */
int gps_ceiling_scratch_bytes(int kind, uint64_t bytes, uint64_t* scratch);

/* One launch of probe `kind` on `stream`: copies `bytes` from src to dst
 * (copy kinds and MEMCPY), reads `bytes` of src (READ; dst receives at most one
 * word, so it needs 4 bytes), or writes `bytes` to dst (WRITE and MEMSET; src
 * unused). src and dst are 16-byte aligned device buffers that do not
 * overlap. The ticketed kinds other than the uncached-status ones need
 * `scratch`: 256-byte aligned, at least gps_ceiling_scratch_bytes, zero
 * before the first launch; each launch leaves its ticket word zero again.
 * The uncached-status kinds ignore `scratch` and use a block of uncached
 * device memory per device that the library allocates on first use (or
 * when a larger run needs more) and never frees. */
int gps_ceiling_run(int kind, const void* src, void* dst, uint64_t bytes, void* scratch, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GPS_CEILING_H */
