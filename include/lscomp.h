/*
 * lscomp.h -- C ABI of the MI355X (gfx950) lsCOMP uint16 codec.
 *
 * Drop-in boundary for the reference's uint16 entry points
 *   void cuLSZ_compression_uint16_bsize64(uint16_t* d_oriData, unsigned char* d_cmpBytes,
 *        size_t* cmpSize, uint3 dims, uint4 quantBins, float poolingSH, cudaStream_t stream=0);
 *   void cuLSZ_decompression_uint16_bsize64(uint16_t* d_decData, unsigned char* d_cmpBytes,
 *        size_t cmpSize, uint3 dims, uint4 quantBins, float poolingSH, cudaStream_t stream=0);
 * (lsCOMP/include/cuLSZ_entry.h:10-11, lsCOMP/src/cuLSZ_entry.cu:93-166).  The
 * same-named C++ shims live in lscomp_culsz.h.
 *
 * Volume layout: dims.x frames (slowest) of dims.y rows of dims.z pixels
 * (fastest), uint16, as in lsCOMP_xpcs.cpp:26.  Each frame is cut into 8x8
 * blocks (zero padded at frame edges); every block gets a 1-byte header and
 * a bit-plane payload.  Stream = [header of every block][zero pad to a multiple
 * of 4][payloads in block order].  Dims and bins are not stored in the stream:
 * decompression needs the same dims and bins; pooling_th only affects
 * compression.
 *
 * Lossless: bins {1,1,1,1} and pooling_th = 1.  Lossy: bins x<=y<=z<=w > 0,
 * pooling_th in [0,1] (a block is pooled when zero_count/64 > pooling_th).
 *
 * All pointers with a d_ prefix are device pointers on the current device.
 * Calls are stream-ordered on `stream` (0 = null stream); the library keeps a
 * small look-back scratch area per (device, stream), allocated on first use.
 * Calls that overlap on one stream are ordered by the stream; calls on
 * different streams run concurrently with separate scratch.
 */
#ifndef LSCOMP_H
#define LSCOMP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lscomp_dim3 { uint32_t x, y, z; } lscomp_dim3;          /* uint3 of the reference */
typedef struct lscomp_bins4 { uint32_t x, y, z, w; } lscomp_bins4;     /* uint4 of the reference */

enum {
    LSCOMP_OK = 0,
    LSCOMP_ERR_INVALID_ARG = 1,  /* dims/bins/pooling out of domain, null or misaligned pointer */
    LSCOMP_ERR_CAPACITY = 2,     /* output buffer too small; *out_bytes holds the required size */
    LSCOMP_ERR_CORRUPT = 3,      /* stream shorter than its headers say, or a header rate > 16 */
    LSCOMP_ERR_HIP = 4,          /* a HIP runtime call failed */
    LSCOMP_ERR_TOO_LARGE = 5,    /* more than 2^31-1 blocks in one volume */
    LSCOMP_ERR_INTERNAL = 6,     /* look-back did not resolve (should not happen) */
    LSCOMP_ERR_RANGE = 7         /* uint32: a block's quantized maximum needs 32 bits (rate 32) */
};

/* Number of 8x8 blocks, x*ceil(y/8)*ceil(z/8) (cuLSZ_entry.cu:97-99). */
size_t lscomp_num_blocks_u16(lscomp_dim3 dims);
/* Header section size R = 4*ceil(blocks/4) (cuLSZ_kernel.cu:843). */
size_t lscomp_header_bytes_u16(lscomp_dim3 dims);
/* Worst-case stream size R + 128*blocks.  A buffer this large never overflows;
 * the reference sizes it nbEle*2 (lsCOMP_xpcs.cpp:137), which can be short. */
size_t lscomp_max_compressed_bytes_u16(lscomp_dim3 dims);

/* Compress; replaces cuLSZ_compression_uint16_bsize64 (cuLSZ_entry.h:10).
 * Blocks the host until the stream size is known and stores it in *out_bytes
 * (R + sum of payloads; the reference reports sum + blocks, cuLSZ_entry.cu:125).
 * If out_capacity is short, nothing past it is written, *out_bytes holds the
 * required size and LSCOMP_ERR_CAPACITY is returned.  d_in and d_out must be
 * 2- and 4-byte aligned; 16-byte aligned d_in with dims.z % 8 == 0 is fastest. */
int lscomp_compress_u16(const uint16_t *d_in, uint8_t *d_out, size_t out_capacity,
                        size_t *out_bytes, lscomp_dim3 dims, lscomp_bins4 bins,
                        float pooling_th, hipStream_t stream);

/* Stream-ordered compress: no host synchronization.  The stream size is
 * written to device memory *d_out_bytes (8 bytes) when the kernel finishes;
 * a size above out_capacity means the payload was not written. */
int lscomp_compress_u16_async(const uint16_t *d_in, uint8_t *d_out, size_t out_capacity,
                              uint64_t *d_out_bytes, lscomp_dim3 dims, lscomp_bins4 bins,
                              float pooling_th, hipStream_t stream);

/* Size of the per-GPU device array behind lscomp_compress_u16_noalloc_async:
 * 8 KiB of control block, then one 8-byte status word per tile. */
#define LSCOMP_NOALLOC_SCRATCH_BYTES (4u << 20)

/* lscomp_compress_u16_async that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, stream and result as lscomp_compress_u16_async; the
 * look-back scratch is a fixed device array per GPU (the library's only
 * allocation is that array, made when the library is loaded and held until
 * the process ends).  Limit: ceil(blocks / tile) <= (4 MiB - 8 KiB) / 8 =
 * 523264 tiles at the current compression tile (128, 256 or 512 blocks), else
 * LSCOMP_ERR_TOO_LARGE.  Calls that overlap on one GPU (unordered streams)
 * race on the array: order them.  lscomp_stream_status does not see this
 * call's device error word; the only error it could hold for compression is
 * a look-back time-out, an internal fault. */
int lscomp_compress_u16_noalloc_async(const uint16_t *d_in, uint8_t *d_out, size_t out_capacity,
                                      uint64_t *d_out_bytes, lscomp_dim3 dims, lscomp_bins4 bins,
                                      float pooling_th, hipStream_t stream);

/* Decompress; replaces cuLSZ_decompression_uint16_bsize64 (cuLSZ_entry.h:11).
 * Waits for completion and returns LSCOMP_ERR_CORRUPT if the stream is
 * inconsistent (those blocks decode to zeros).  Every element of the volume is
 * written, rate-0 blocks as zeros. */
int lscomp_decompress_u16(uint16_t *d_out, const uint8_t *d_in, size_t in_bytes,
                          lscomp_dim3 dims, lscomp_bins4 bins, hipStream_t stream);

/* Stream-ordered decompress: no host synchronization; stream errors are
 * sticky and reported by lscomp_stream_status(). */
int lscomp_decompress_u16_async(uint16_t *d_out, const uint8_t *d_in, size_t in_bytes,
                                lscomp_dim3 dims, lscomp_bins4 bins, hipStream_t stream);

/* Synchronizes `stream`, returns and clears errors raised by earlier async
 * calls on it (LSCOMP_ERR_CORRUPT / LSCOMP_ERR_INTERNAL), or LSCOMP_OK. */
int lscomp_stream_status(hipStream_t stream);

/* Pre-allocates the scratch of `stream` for volumes up to `dims` so that
 * later calls on it perform no allocation (e.g. before hipGraph capture). */
int lscomp_reserve_u16(lscomp_dim3 dims, hipStream_t stream);

/* Releases every scratch area of the current device (synchronizes). */
void lscomp_release_all(void);

/* Launch shape of later calls (process-wide): direction 0 = compression,
 * 1 = decompression; tile = 8x8 blocks per workgroup (128, 256 or 512;
 * 0 = the default, 512); wg_per_cu =
 * workgroups per CU (0 = occupancy limit).
 * Streams are identical for every shape; only speed changes.  No counterpart
 * in the reference (its grid is fixed, cuLSZ_entry.cu:100-101); set by the
 * autotuner or by hand. */
int lscomp_set_launch_shape(int direction, uint32_t tile, uint32_t wg_per_cu);
int lscomp_get_launch_shape(int direction, uint32_t *tile, uint32_t *wg_per_cu);

/* The full launch shape.  lb_window: look-back words read per probe round
 * (a multiple of 64, at most tile; 0 = the default, 64).  counters: device
 * counters that hand out tile ids (1..16; 0 = the default, 4 for compression
 * and 2 for decompression): one same-address atomic per tile serializes near
 * 80 tiles per microsecond, so tile ids are spread over several counters, each
 * workgroup drawing from one.
 * lscomp_set_launch_shape sets tile and wg_per_cu and the default window and
 * counters.  Streams are identical for every shape.  No counterpart in the
 * reference (cuLSZ_entry.cu:100-101, cuLSZ_kernel.h:7 fix its grid). */
typedef struct lscomp_launch_shape {
    uint32_t tile, wg_per_cu, lb_window, counters;
} lscomp_launch_shape;
int lscomp_set_launch_shape_ex(int direction, const lscomp_launch_shape *shape);
int lscomp_get_launch_shape_ex(int direction, lscomp_launch_shape *shape);
/* How single-volume decompression finds each tile's payload offset:
 * 1 = from the header bytes in a first phase of the decode launch (default),
 * 2 = from the header bytes in a separate pre-pass kernel, 0 = a decoupled
 * look-back inside the decode kernel.  Returns the previous mode (negative
 * for an invalid mode).  Output is identical in every mode. */
int lscomp_set_decode_prepass(int mode);

/* One measured launch shape (ms per call, averaged over the repetitions). */
typedef struct lscomp_tune_entry {
    uint32_t direction;        /* 0 compression, 1 decompression */
    uint32_t tile, wg_per_cu, lb_window, counters;
    float ms;
} lscomp_tune_entry;

/* Autotuner: times compression and then decompression of the caller's volume
 * (d_in, dims, bins, pooling) under launch shapes found by coordinate descent
 * over tile (128/256/512), workgroups per CU (each value giving a distinct
 * grid), look-back window and tile-id counters, from the current shape of
 * each direction, using d_stream (capacity >= lscomp_max_compressed_bytes_u16
 * (dims)) and d_decoded (same size as the volume) as work buffers; applies the
 * fastest shape of each direction with lscomp_set_launch_shape_ex.  Results
 * are cached per (size class, mode): a later call for the same class applies
 * the cached shapes without timing.  If table/n_table are given, the measured
 * entries are copied there (up to *n_table; *n_table is set to the number
 * written, 0 for a cache hit).  No counterpart in the reference. */
int lscomp_autotune_u16(const uint16_t *d_in, uint8_t *d_stream, size_t stream_capacity, uint16_t *d_decoded,
                        lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th, int reps, hipStream_t stream,
                        lscomp_tune_entry *table, uint32_t *n_table);

/* ---- Batched form -------------------------------------------------------
 * Many independent volumes (e.g. detector frames of uneven sizes), each with
 * its own stream in exactly the plain format, processed by ONE persistent
 * kernel launch that pulls tiles of every item from a shared queue.  No
 * counterpart in the reference (one volume per launch there); a plan's item
 * streams are byte-identical to lscomp_compress_u16 on each item alone.
 *
 * A plan owns its device tables and look-back scratch (allocated once by
 * lscomp_plan_create); lscomp_plan_set_items_u16 refills it without
 * allocating, so one plan serves batch after batch.  Use one stream per plan. */
typedef struct lscomp_item_u16 {
    lscomp_dim3 dims;          /* x frames of y rows of z pixels */
    lscomp_bins4 bins;
    float pooling_th;          /* compression only */
    uint16_t *d_volume;        /* compress: input volume; decompress: output volume */
    uint8_t *d_stream;         /* compress: output stream; decompress: input stream */
    uint64_t stream_bytes;     /* compress: capacity of d_stream; decompress: stream length */
} lscomp_item_u16;

typedef struct lscomp_plan_s *lscomp_plan;

/* direction 0 = compress, 1 = decompress.  Capacity: up to max_items items and
 * max_blocks 8x8 blocks in total per batch.  Starts with the launch shape of
 * the direction at creation time (lscomp_set_launch_shape[_ex]). */
int lscomp_plan_create(int direction, uint32_t max_items, uint64_t max_blocks, lscomp_plan *plan);
/* Validates the items and uploads the batch tables (stream-ordered on `stream`;
 * the host array may be reused when the call returns). */
int lscomp_plan_set_items_u16(lscomp_plan plan, const lscomp_item_u16 *items, uint32_t n, hipStream_t stream);
/* The plan's launch shape (tile, workgroups per CU, look-back window,
 * counters; zeros take the direction's defaults).  Rebuilds the tables of the
 * current batch when the tile changes and drops a captured graph (the next
 * lscomp_plan_launch_graph captures again).  Streams are identical for every
 * shape.  A plan starts with the direction's shape at creation. */
int lscomp_plan_set_shape(lscomp_plan plan, const lscomp_launch_shape *shape, hipStream_t stream);
int lscomp_plan_get_shape(lscomp_plan plan, lscomp_launch_shape *shape);
/* Autotunes the plan's shape on its current batch (the same search as
 * lscomp_autotune_u16; each measured launch rewrites the batch's outputs with
 * the same bytes) and applies the fastest.  table / n_table as there. */
int lscomp_plan_autotune(lscomp_plan plan, int reps, hipStream_t stream, lscomp_tune_entry *table,
                         uint32_t *n_table);
/* One persistent launch over every item of the current batch (stream-ordered). */
int lscomp_plan_launch(lscomp_plan plan, hipStream_t stream);
/* The same launch replayed from a hipGraph captured on first use (batches set
 * later are picked up: the graph reads the tables through pointers). */
int lscomp_plan_launch_graph(lscomp_plan plan, hipStream_t stream);
/* Device array of per-item stream sizes (compress) written by the launch. */
const uint64_t *lscomp_plan_device_sizes(lscomp_plan plan);
/* Synchronizes `stream` and copies the per-item stream sizes to the host. */
int lscomp_plan_sizes(lscomp_plan plan, uint64_t *host_sizes, hipStream_t stream);
/* Synchronizes `stream`; returns and clears errors of the plan's launches. */
int lscomp_plan_status(lscomp_plan plan, hipStream_t stream);
int lscomp_plan_destroy(lscomp_plan plan);

/* ---- Served form --------------------------------------------------------
 * A compression server with continuous batching.  Requests (device frames)
 * are queued by lscomp_server_submit_u16 and admitted in arrival order while
 * their worst-case output fits into an HBM arena of `budget_bytes`; each
 * request's stream is exactly lscomp_compress_u16 of its frame alone, and its
 * output range stays reserved until lscomp_server_release.  Two engines:
 *  - lscomp_server_create_device: a device-resident request loop.  One
 *    persistent kernel takes requests from a ring in host-pinned memory as
 *    they are admitted, spreads their tiles over every workgroup (requests
 *    that arrive while others run join at once) and writes each completion
 *    into host-pinned memory; no kernel launch per request or batch.  A
 *    kernel instance ends after `idle_us` without a request (or when its
 *    tables are full) and a host thread starts the next one when requests
 *    wait.  max_blocks bounds one request.
 *  - lscomp_server_create: a host thread packs the requests admitted since its
 *    last launch into one launch of one of `slots` plans (batched form), each
 *    on its own stream, so batches overlap; max_items / max_blocks per batch.
 * No counterpart in the reference. */
typedef struct lscomp_request_u16 {
    lscomp_dim3 dims;
    lscomp_bins4 bins;
    float pooling_th;
    const uint16_t *d_frame;   /* device input, alive and unchanged until the request completes */
    uint64_t tag;              /* returned with the completion */
} lscomp_request_u16;

typedef struct lscomp_completion {
    uint64_t id;               /* from lscomp_server_submit_u16 */
    uint64_t tag;
    int status;                /* LSCOMP_OK, or an error for this request (LSCOMP_ERR_INTERNAL: the
                                  launch that ran it reported a look-back timeout) */
    uint32_t reserved;
    uint64_t stream_bytes;
    const uint8_t *d_stream;   /* in the server's arena; valid until released (null on error) */
    double t_submit, t_admit, t_launch, t_done;   /* seconds (steady clock) */
} lscomp_completion;

typedef struct lscomp_server_s *lscomp_server;

typedef struct lscomp_server_stats_t {
    uint64_t submitted, completed;
    uint64_t launches;         /* kernel launches: batches (host engine) or loop instances (device engine) */
    uint32_t device_loop;      /* 1 for the device engine */
    uint32_t reserved;
} lscomp_server_stats_t;

/* Host engine: slots = batches in flight (streams); max_items / max_blocks per batch. */
int lscomp_server_create(uint32_t slots, uint32_t max_items, uint64_t max_blocks, uint64_t budget_bytes,
                         lscomp_server *srv);
/* Device engine: max_blocks per request, idle_us before an idle kernel instance ends. */
int lscomp_server_create_device(uint64_t max_blocks, uint64_t budget_bytes, uint32_t idle_us, lscomp_server *srv);
/* LSCOMP_ERR_TOO_LARGE for more than max_blocks blocks, LSCOMP_ERR_CAPACITY
 * when the request's worst-case stream exceeds the whole budget. */
int lscomp_server_submit_u16(lscomp_server srv, const lscomp_request_u16 *req, uint64_t *id);
/* Non-blocking: copies up to max finished requests to out, sets *n. */
int lscomp_server_poll(lscomp_server srv, lscomp_completion *out, uint32_t max, uint32_t *n);
/* Frees the output buffer of a finished request (makes room for admission). */
int lscomp_server_release(lscomp_server srv, uint64_t id);
/* Blocks until every submitted request has finished; returns the first error
 * a completion reported since the last drain (or LSCOMP_OK). */
int lscomp_server_drain(lscomp_server srv);
int lscomp_server_stats(lscomp_server srv, lscomp_server_stats_t *out);
/* Submits reqs[i] at time t0 + arrival_s[i] (busy-waiting on the calling
 * thread, t0 = call time), collects every completion into out[i] (by request
 * order); returns when all finished.  By default each output is released as
 * soon as it is collected (d_stream is then null); with LSCOMP_REPLAY_HOLD
 * the outputs stay reserved until the caller releases them.  A trace replay
 * for load tests (e.g. Poisson arrivals). */
enum { LSCOMP_REPLAY_HOLD = 1 };
int lscomp_server_replay_u16(lscomp_server srv, const lscomp_request_u16 *reqs, const double *arrival_s,
                             uint32_t n, lscomp_completion *out, uint32_t flags);
int lscomp_server_destroy(lscomp_server srv);

/* ---- Error metrics ------------------------------------------------------
 * Quality of a lossy round trip, computed on the GPU in one pass over both
 * volumes; mse, range, psnr, norm_err, norm_err_norm and nrmse are the six
 * results of the reference's computePSNR (lsCOMP/src/cuLSZ_utility.cu:887-946),
 * with exact integer sums for uint16 data (valid while the sums fit 64 bits). */
typedef struct lscomp_error_stats {
    uint64_t n;
    uint32_t max_abs_err;          /* max |dec - orig| */
    uint32_t min_orig, max_orig;
    uint32_t reserved;
    uint64_t sum_sq_err, sum_sq_dec;
    double max_rel_err;            /* max |dec - orig| / orig over orig != 0 */
    double mse, range, psnr, norm_err, norm_err_norm, nrmse;
} lscomp_error_stats;

int lscomp_error_stats_u16(const uint16_t *d_orig, const uint16_t *d_dec, size_t n, lscomp_error_stats *out,
                           hipStream_t stream);

/* Mean SSIM of two uint16 volumes (x, y, z), z fastest, into *ssim: replaces
 * the reference's computeSSIM(ori, dec, x, y, z) (lsCOMP/src/cuLSZ_utility.cu:
 * 844-886, windows 757-842): 7x7x7 windows moved by 2 along every axis,
 * luminance * contrast * structure with c1 = (0.01 R)^2, c2 = (0.03 R)^2 from
 * the original's range R in the window (0.01^2, 0.03^2 when R = 0), averaged
 * over the windows.  Window moments are exact (the reference rounds two-pass
 * double sums), so results agree to about 1e-12 relative.  Where the
 * reference prints an error and returns 0/0 (a dimension below 7), this
 * returns LSCOMP_ERR_INVALID_ARG.  Synchronous on `stream`; allocates one
 * double per workgroup. */
int lscomp_ssim_u16(const uint16_t *d_orig, const uint16_t *d_dec, lscomp_dim3 dims, double *ssim,
                    hipStream_t stream);

/* ---- uint32 volumes --------------------------------------------------------
 * Replace cuLSZ_compression_uint32_bsize64 / cuLSZ_decompression_uint32_bsize64
 * (lsCOMP/include/cuLSZ_entry.h:8-9, lsCOMP/src/cuLSZ_entry.cu:18-91): same
 * blocking, decision and stream format as uint16 with 32-bit values (the pair
 * average wraps modulo 2^32 like the reference's uint sum, rates up to 31).
 * A block whose quantized maximum needs 32 bits cannot be represented (the
 * reference corrupts its header and leaves an unwritten gap,
 * cuLSZ_kernel.cu:120-131, 259-261): compression returns LSCOMP_ERR_RANGE
 * (use larger bins).  Same semantics as the uint16 entries otherwise; the
 * volume and stream pointers must be 4-byte aligned.  Worst case stream:
 * lscomp_max_compressed_bytes_u32 = R + 248 per block.  The header and block
 * counts are those of lscomp_header_bytes_u16 / lscomp_num_blocks_u16. */
size_t lscomp_max_compressed_bytes_u32(lscomp_dim3 dims);
int lscomp_compress_u32(const uint32_t *d_in, uint8_t *d_out, size_t out_capacity, size_t *out_bytes,
                        lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th, hipStream_t stream);
int lscomp_compress_u32_async(const uint32_t *d_in, uint8_t *d_out, size_t out_capacity, uint64_t *d_out_bytes,
                              lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th, hipStream_t stream);
int lscomp_decompress_u32(uint32_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                          lscomp_bins4 bins, hipStream_t stream);
int lscomp_decompress_u32_async(uint32_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                                lscomp_bins4 bins, hipStream_t stream);

/* ---- self-describing container --------------------------------------------
 * The reference passes dims, bins and stream sizes out of band
 * (cuLSZ_entry.h:8-11).  A container stores any number of streams (a volume,
 * or the frames of a batch or of the served form) with their parameters:
 * a 48-byte header ("LSCP", version 1, item count, total bytes), a 64-byte
 * little-endian index entry per item, then the streams at 16-byte-aligned
 * offsets.  Host memory only (copy device streams in and out); no
 * counterpart in the reference. */
typedef struct lscomp_container_item {
    lscomp_dim3 dims;
    uint32_t elem_bytes;       /* 2: uint16 volume, 4: uint32 volume */
    lscomp_bins4 bins;
    float pooling_th;          /* as used for compression (informational) */
    uint32_t reserved;
    uint64_t offset;           /* of the stream in the container (set by pack and parse) */
    uint64_t stream_bytes;
    uint64_t tag;              /* caller's (e.g. frame number) */
} lscomp_container_item;

/* Bytes of a container holding these items (their stream_bytes). */
size_t lscomp_container_bytes(const lscomp_container_item *items, uint32_t n_items);
/* Writes the container to out (capacity bytes; *out_bytes = its size) from
 * host streams[i] of items[i].stream_bytes bytes; sets items[i].offset.
 * LSCOMP_ERR_INVALID_ARG for parameters the codec would reject or a stream
 * size outside [header bytes, worst case]; LSCOMP_ERR_CAPACITY if too small. */
int lscomp_container_pack(lscomp_container_item *items, uint32_t n_items, const uint8_t *const *streams,
                          uint8_t *out, size_t capacity, size_t *out_bytes);
/* Reads the index of a container of len bytes into items (up to max_items;
 * *n_items = the container's count).  Stream i is buf + items[i].offset.
 * LSCOMP_ERR_CORRUPT for a malformed container (magic, version, sizes,
 * offsets out of order or range, invalid parameters), LSCOMP_ERR_CAPACITY if
 * max_items is too small. */
int lscomp_container_parse(const uint8_t *buf, size_t len, lscomp_container_item *items, uint32_t max_items,
                           uint32_t *n_items);

/* Measurement helper: a plain device copy of `bytes` bytes (d_dst, d_src
 * 16-byte aligned, non-overlapping), one 16-byte load and store per lane like
 * the codec kernels' row accesses.  Gives the copy bandwidth and the scale of
 * the FETCH_SIZE / WRITE_SIZE counters for such accesses.  No counterpart in
 * the reference. */
int lscomp_copy_bytes(void *d_dst, const void *d_src, size_t bytes, hipStream_t stream);

const char *lscomp_status_string(int status);

/* Library version, major*10000 + minor*100 + patch. */
int lscomp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LSCOMP_H */
