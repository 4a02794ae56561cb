/* zero_hip.h: the C ABI of the MI355X (gfx950) zero block codec, libzero_hip.so.
 *
 * It replaces the reference's GPU runtime layer (src/gpu.h, src/ze_min.h: device, buffers, kernel launches) and the
 * kernel sequences that src/gpu_codec.h drives, with ahead-of-time compiled HIP kernels. Plain pointers and sizes
 * only; every function returns ZGPU_OK (0) or a negative ZGPU_ERR_* code, with a message from zgpu_last_error().
 * A context is used by one host thread at a time (as the reference's GpuCodec is).
 *
 * Formats (bit-exact with the reference): the ZB2 block payload (src/codec2.h:16-17,140-151) with its rANS stream
 * sections (src/codec.h:4-6,96-147), and the CRC-32 of src/crc32.h.
 */
#ifndef ZERO_HIP_H
#define ZERO_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZGPU_OK 0
#define ZGPU_ERR_DEVICE (-1)  /* the HIP runtime failed (launch failure, lost device): src/gpu.h:17-27 gpu::Error */
#define ZGPU_ERR_ARG (-2)     /* an argument is invalid (block size, level, sizes) */
#define ZGPU_ERR_DAMAGED (-3) /* at least one block is damaged (per-block bits in err[]): src/codec.h:29-31 zb::Error */
#define ZGPU_ERR_NOMEM (-4)   /* device memory exhausted */

/* per-block error bits written to err[] by the decoder */
#define ZGPU_E_HDR 0x01u
#define ZGPU_E_COUNTS 0x02u
#define ZGPU_E_STREAM 0x04u
#define ZGPU_E_TABLE 0x08u
#define ZGPU_E_LANE 0x10u
#define ZGPU_E_TOKEN 0x20u
#define ZGPU_E_GROUP 0x40u
#define ZGPU_E_METHOD 0x80u

#define ZGPU_STORED 0xFFFFFFFFu /* out_len of a block whose payload would not be smaller: write it stored */
#define ZGPU_BAD_BLOCK 0xFFFFFFFEu /* out_len of a block table entry with raw_len > 2^block_log: nothing encoded */

typedef struct zgpu zgpu_t;
typedef void* zgpu_stream_t; /* a hipStream_t; NULL is the context's own stream */

/* One entry per block of a launch (32 bytes, device memory for the zgpu_*_blocks calls). Encoding reads raw_len bytes
 * at in + src, at `level` (1..9) when it is not 0, else at the call's level. Decoding reads comp_len bytes at
 * comp + src (method 0 stored, 1 ZB2) and writes raw_len bytes at out + dst. Mirrors zero::BlockIn
 * (src/archive.h:50-55) plus where the bytes live. */
typedef struct {
  uint64_t src;
  uint64_t dst;
  uint32_t raw_len;
  uint32_t comp_len;
  uint32_t method;
  uint32_t level;
} zgpu_block_t;

typedef struct {
  char name[128];      /* marketing name */
  char arch[64];       /* gcnArchName, e.g. gfx950 */
  int cu_count;
  int clock_khz;
  uint64_t hbm_bytes;
  uint32_t lds_per_block; /* LDS a workgroup may use */
  uint32_t max_group;
  int integrated;
} zgpu_props_t;

/* ---- device (src/gpu.h:124-257 gpu::Device::open / sync / properties) */
int zgpu_open(int ordinal, zgpu_t** out);
void zgpu_close(zgpu_t* g);
int zgpu_props(zgpu_t* g, zgpu_props_t* out);
const char* zgpu_last_error(zgpu_t* g);
int zgpu_sync(zgpu_t* g);

/* ---- memory (src/gpu.h:200-252 gpu::Buffer alloc / alloc_host / write / read / zero) */
int zgpu_alloc(zgpu_t* g, size_t bytes, void** dptr);
int zgpu_free(zgpu_t* g, void* dptr);
int zgpu_alloc_host(zgpu_t* g, size_t bytes, void** hptr); /* pinned */
int zgpu_free_host(zgpu_t* g, void* hptr);
int zgpu_upload(zgpu_t* g, void* dptr, const void* host, size_t bytes);   /* synchronous */
int zgpu_download(zgpu_t* g, void* host, const void* dptr, size_t bytes); /* synchronous */
int zgpu_fill(zgpu_t* g, void* dptr, uint8_t value, size_t bytes, zgpu_stream_t s); /* a kernel, graph-safe */
/* device-to-device copy on the stream by a plain 16-byte-per-lane kernel (k_copy) when both pointers are 16-byte
   aligned: the codec's measured bandwidth ceiling and the calibration kernel for HBM byte counters */
int zgpu_copy(zgpu_t* g, void* dst, const void* src, size_t bytes, zgpu_stream_t s);

/* ---- streams and events (timing on the stream the kernels run on) */
int zgpu_stream_create(zgpu_t* g, zgpu_stream_t* out);
int zgpu_stream_destroy(zgpu_t* g, zgpu_stream_t s);
int zgpu_stream_sync(zgpu_t* g, zgpu_stream_t s);
int zgpu_event_create(zgpu_t* g, void** ev);
int zgpu_event_destroy(zgpu_t* g, void* ev);
int zgpu_event_record(zgpu_t* g, void* ev, zgpu_stream_t s);
int zgpu_event_elapsed_ms(zgpu_t* g, void* start, void* stop, float* ms);

/* ---- the codec on device-resident buffers (src/gpu_codec.h:153-229 compress, :253-315 decompress)
 * block_log 16..20; level 1..9 (src/gpu_codec.h:66-70: 1-2 two neighbours / 16 bytes compared, 3-6 four / 16,
 * 7-9 eight / 32). Scratch is the context's, grown on demand. */

/* Encodes nb blocks (each raw_len <= 2^block_log). Block b's payload goes to out + (b << block_log) and its size to
 * out_len[b], or ZGPU_STORED when it would not be smaller than the block. The kernels check the table themselves: an
 * entry with raw_len > 2^block_log is read as an empty block and gets out_len ZGPU_BAD_BLOCK, and nothing is read
 * or written outside its slab. More than 2^31 bytes of blocks (8192 of 2^18) run as several launches in turn, over
 * scratch for one of them. */
int zgpu_encode_blocks(zgpu_t* g, const uint8_t* d_in, const zgpu_block_t* d_blocks, uint32_t nb, uint32_t block_log,
                       int level, uint8_t* d_out, uint32_t* d_out_len, zgpu_stream_t s);
/* Decodes nb blocks into out; err[b] receives ZGPU_E_* bits (0 = good). Asynchronous: read err after syncing. The
 * kernels check the table themselves: raw_len > 2^block_log sets ZGPU_E_COUNTS before any byte of the block is read
 * or written, and every other length comes from the payload and is checked before use (damaged input never makes
 * the kernels access memory outside the block's payload, its scratch and its raw_len bytes at out + dst). */
int zgpu_decode_blocks(zgpu_t* g, const uint8_t* d_comp, const zgpu_block_t* d_blocks, uint32_t nb, uint32_t block_log,
                       uint8_t* d_out, uint32_t* d_err, zgpu_stream_t s);
/* zgpu_decode_blocks without any allocation. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The decoder's
 * scratch lives in a fixed array of ZGPU_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on until
 * the process ends: 128 blocks of 2^18, 32 of 2^20 or 512 of 2^16 per call (zgpu_decode_scratch_bytes(nb,
 * block_log) at most that); a call that needs more is refused with ZGPU_ERR_ARG. The call allocates nothing and
 * records no event. Every call on a GPU shares that GPU's array, which keeps the last call's contents: calls that
 * overlap on one GPU (other threads, other streams, other contexts) race, so a caller runs them in turn. */
#define ZGPU_NOALLOC_SCRATCH_BYTES (64ull << 20)
int zgpu_decode_blocks_noalloc(zgpu_t* g, const uint8_t* d_comp, const zgpu_block_t* d_blocks, uint32_t nb,
                               uint32_t block_log, uint8_t* d_out, uint32_t* d_err, zgpu_stream_t s);
/* crc[b] = CRC-32 of block b's raw bytes, at data + dst (use_dst != 0) or data + src (src/archive.h:376,618) */
int zgpu_crc32_blocks(zgpu_t* g, const uint8_t* d_data, const zgpu_block_t* d_blocks, uint32_t nb, int use_dst,
                      uint32_t* d_crc, zgpu_stream_t s);
/* device scratch the context holds for nb blocks of 2^block_log */
size_t zgpu_encode_scratch_bytes(uint32_t nb, uint32_t block_log);
size_t zgpu_decode_scratch_bytes(uint32_t nb, uint32_t block_log);

/* ---- the codec on host buffers: zero::Codec::compress / decompress (src/archive.h:56-69) as GpuCodec does them.
 * Synchronous. A batch moves in chunks over two streams, so the PCIe copies of one chunk overlap the kernels of the
 * other (the reference uploads, runs and downloads in turn, src/gpu_codec.h:153-229,253-315); host buffers
 * allocated with zgpu_alloc_host (pinned) move fastest and overlap best. */
/* n bytes = whole blocks, the last may be shorter. out: the payloads of the compressed blocks back to back, in block
 * order (a stored block takes no room; n bytes always suffice); out_len[b] = block b's payload size, or
 * ZGPU_STORED. crc (may be NULL): the CRC-32 of each block's raw bytes, which the .zero container stores per block
 * (src/archive.h:376). */
int zgpu_compress(zgpu_t* g, const uint8_t* in, uint64_t n, uint32_t block_log, int level, uint8_t* out,
                  uint64_t out_cap, uint32_t* out_len, uint32_t* crc);
/* blocks back to back in comp (comp_len[b] bytes each), their raw bytes back to back into out. err[b] as above;
 * returns ZGPU_ERR_DAMAGED when any block is damaged. crc (may be NULL): the CRC-32 of each decoded block, for the
 * container's check (src/archive.h:618). */
int zgpu_decompress(zgpu_t* g, const uint8_t* comp, const uint32_t* comp_len, const uint8_t* method,
                    const uint32_t* raw_len, uint32_t nb, uint32_t block_log, uint8_t* out, uint32_t* err,
                    uint32_t* crc);
/* One rANS stream section of n symbols (host buffers) by the block encoder's own kernels: histogram, 12-bit
 * normalisation with its tie rules, 64 interleaved lanes, raw when not smaller (src/codec.h:41-147
 * zb::normalize + zb::encode_stream, which it replaces). out needs n + 5 + 420 bytes; *out_len = the section's size.
 * n <= 2^19. */
int zgpu_encode_stream(zgpu_t* g, const uint8_t* sym, uint32_t n, uint8_t* out, uint32_t* out_len);
/* crc[b] of the nb blocks of 2^block_log in data[0..n) (host buffer) */
int zgpu_crc32(zgpu_t* g, const uint8_t* data, uint64_t n, uint32_t block_log, uint32_t* crc);
/* crc[i] = CRC-32 of bytes [off[i], off[i] + len[i]) of the last zgpu_compress input (which = 0) or the last
 * zgpu_decompress output (which = 1), which stay on the device until the next host-buffer call on g. For the
 * container's per-file checksums of the solid stream, which the reference computes on the host while it reads and
 * writes files (src/archive.h:279-287,320,707,774). ZGPU_ERR_ARG for a range outside that data, or when another
 * host-buffer call came in between. */
int zgpu_crc32_ranges(zgpu_t* g, int which, uint32_t n, const uint64_t* off, const uint64_t* len, uint32_t* crc);

/* ---- the batched form: the blocks of many files, of uneven lengths, planned together (no reference counterpart:
 * the reference launches its stages per 16 MiB batch of full blocks, src/gpu_codec.h:153-229, and assumes every
 * block but the last is full, :160). A block of r bytes runs in a slab of 2^c bytes, c = max(10, ceil(log2 r)) up
 * to block_log (its size class). Every stage of the encoder and decoder is one persistent launch over the blocks of
 * every class, whose workgroups take work items from per-XCD device queues, largest class first; blocks run in
 * passes of at most 2 GiB of slabs (one scratch arena, reused). Every payload is byte-identical to the plain
 * path's (the hash keys keep block_log's width). */
typedef struct zgpu_plan zgpu_plan_t;
/* nb blocks: block b's raw bytes are raw_len[b] (<= 2^block_log) bytes at offset src[b] of the input buffer (host
 * arrays, copied). streams 1..8 (kept for callers: the persistent stages spread over the whole GPU on one). */
int zgpu_plan_create(zgpu_t* g, uint32_t nb, const uint64_t* src, const uint32_t* raw_len, uint32_t block_log,
                     int streams, zgpu_plan_t** out);
void zgpu_plan_destroy(zgpu_plan_t* p);
/* bytes of the output slabs zgpu_plan_encode writes; slab_off[b] (host, nb) = where block b's payload goes */
uint64_t zgpu_plan_slab_bytes(const zgpu_plan_t* p);
int zgpu_plan_slab_offsets(const zgpu_plan_t* p, uint64_t* slab_off);
/* the size classes: per class its slab log, block count, bucket bits and parse walkers; returns the class count */
int zgpu_plan_classes(const zgpu_plan_t* p, uint32_t* block_log, uint32_t* blocks, uint32_t* hash_bits,
                      uint32_t* walkers, int cap);
/* the workgroups per compute unit of each persistent stage (hash, place, sort_s, sort_m, sort_l, match, parse,
 * tokens, rans, pack, decode streams, decode replay); returns the stage count */
int zgpu_plan_stage_occupancy(zgpu_plan_t* p, uint32_t* occ, int cap);
/* sets them (n stages from the first; 1..64 each) */
int zgpu_plan_set_stage_occupancy(zgpu_plan_t* p, const uint32_t* occ, int n);
/* the slab bytes a pass holds at most (2 GiB unless set; at least 1 MiB): the encoder's scratch arena is about 19
 * times that */
int zgpu_plan_pass_bytes(zgpu_plan_t* p, uint64_t slab_bytes);
/* the match stage's threads per workgroup (64, 128, 256, 512 or 1024; 256 unless set or tuned), which are also the
 * sorted slots one of its work items holds; threads = 0 only reads it. Returns the current value. */
int zgpu_plan_match_workgroup(zgpu_plan_t* p, uint32_t threads);
/* on: the next encode or decode records an event after every stage of every pass (graphs are not used meanwhile);
 * zgpu_plan_stage_ms then gives each stage's time summed over the passes (ms[stage], the stage order above) */
int zgpu_plan_profile(zgpu_plan_t* p, int on);
int zgpu_plan_stage_ms(zgpu_plan_t* p, float* ms, int cap);
/* the passes the plan's blocks run in (-1 before the first encode or decode) */
int zgpu_plan_passes(const zgpu_plan_t* p);
/* encodes every block (on the context's stream; asynchronous): payload at d_slabs + slab_off[b], its size (or
 * ZGPU_STORED) at d_out_len[b] */
int zgpu_plan_encode(zgpu_plan_t* p, const uint8_t* d_in, int level, uint8_t* d_slabs, uint32_t* d_out_len);
/* where the stored blocks are for zgpu_plan_decode: block b's comp_len[b] bytes at offset comp_off[b] of the
 * compressed buffer, method[b] 0 stored / 1 ZB2 (host arrays, copied) */
int zgpu_plan_set_stored(zgpu_plan_t* p, const uint64_t* comp_off, const uint32_t* comp_len, const uint8_t* method);
/* decodes every block: raw bytes to d_out + src[b] (the input layout of zgpu_plan_create), err[b] as
 * zgpu_decode_blocks */
int zgpu_plan_decode(zgpu_plan_t* p, const uint8_t* d_comp, uint8_t* d_out, uint32_t* d_err);
/* chooses the output-invariant parameters on this input and keeps the fastest: each class's bucket bits (the
 * hash-bucket depth) and parse walkers (the parse-group width), and each encoder stage's workgroups per compute
 * unit (synchronous; writes d_slabs / d_out_len as zgpu_plan_encode does) */
int zgpu_plan_tune(zgpu_plan_t* p, const uint8_t* d_in, int level, uint8_t* d_slabs, uint32_t* d_out_len);
/* on: a call's whole launch sequence is captured into a hipGraph the first time and replayed when the same buffers
 * and level come again */
int zgpu_plan_use_graphs(zgpu_plan_t* p, int on);

/* ---- the serving loop (no reference counterpart: the reference has no server). A request loop admits compress and
 * decompress requests of mixed sizes and levels into batches by HBM budget while earlier batches run (continuous
 * batching over `slots` batch slots, each with its own stream and preallocated device arenas), and reports latency
 * from arrival to the result in host memory. zgpu_serve_bench drives it with a synthetic open-loop load: Poisson
 * arrivals at offered_GBps of raw bytes for `seconds`, lognormal sizes, compress_share compress requests (level
 * uniform over 1 / 5 / 9), inputs cut from the two pools in random_share : 1 - random_share. A decompress request
 * takes consecutive blocks of a pool as compressed before the clock starts: by the caller (dec_comp / dec_comp_len
 * / dec_method: every 2^block_log block of the pool, back to back; method 0 stored, 1 ZB2), or when those are NULL
 * by the GPU at level 5. After the run every response is round-tripped, and alone_checks requests are served again
 * one per batch and compared with their results in the mix. With sample_every > 0, the responses of requests
 * 0, sample_every, 2 sample_every, ... (in arrival order) are copied out, with what they were asked, until one of
 * the three sample buffers is full. */
typedef struct {
  uint32_t op;          /* 0 compress, 1 decompress */
  uint32_t level;       /* compress: 1..9 */
  uint32_t pool;        /* 0 the LZ pool, 1 the random pool */
  uint32_t nblk;        /* blocks of 2^block_log in the request */
  uint64_t src_off;     /* where its raw bytes start in the pool (a block boundary for decompress) */
  uint64_t raw_len;
  uint64_t bytes_off;   /* its response in sample_bytes: compress, the payloads of its compressed blocks back to
                           back; decompress, its raw bytes */
  uint64_t bytes_len;
  uint32_t blk_len_off; /* compress: its blocks' payload sizes (or ZGPU_STORED) at sample_blk_len[blk_len_off..] */
  uint32_t err;         /* decompress: the ZGPU_E_* bits of its blocks */
  double latency_ms;
} zgpu_serve_sample_t;
typedef struct {
  double offered_GBps, seconds, compress_share, random_share;
  double size_median, size_sigma;
  uint64_t size_max;
  uint64_t hbm_budget;
  uint32_t block_log;
  int slots;
  uint64_t seed;
  uint32_t alone_checks;
  const uint8_t* dec_comp[2];
  const uint32_t* dec_comp_len[2];
  const uint8_t* dec_method[2];
  uint32_t sample_every;
  uint32_t sample_cap;
  zgpu_serve_sample_t* samples;
  uint8_t* sample_bytes;
  uint64_t sample_bytes_cap;
  uint32_t* sample_blk_len;
  uint32_t sample_blk_cap;
} zgpu_serve_cfg_t;
typedef struct {
  uint64_t requests, compress_requests, decompress_requests, bytes;
  double p50_ms, p99_ms, max_ms, mean_ms;
  double wall_s, sustained_GBps;
  uint64_t batches;
  double mean_batch_requests;
  uint64_t hbm_bytes, slot_input_bytes;
  uint64_t mismatches, alone_checked, alone_mismatches;
  uint64_t samples;
} zgpu_serve_report_t;
int zgpu_serve_bench(zgpu_t* g, const zgpu_serve_cfg_t* cfg, const uint8_t* lz_pool, uint64_t lz_len,
                     const uint8_t* random_pool, uint64_t random_len, zgpu_serve_report_t* rep);

/* ---- stage timing (ZERO_PROFILE's per-stage split, src/gpu_codec.h:32-49): when on, events are recorded between
 * the kernels of every encode/decode call on its stream; zgpu_stage_times reads the last call's (it syncs). */
int zgpu_set_profile(zgpu_t* g, int on);
int zgpu_stage_times(zgpu_t* g, const char** names, float* ms, int cap);

#ifdef __cplusplus
}
#endif
#endif
