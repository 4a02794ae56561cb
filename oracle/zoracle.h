/* The oracle: a CPU restatement of zero's block codec in plain C.
 *
 * TEST INFRASTRUCTURE ONLY. It is the checker that the HIP path is compared against; nothing in the
 * product library, CLI or benchmark timing links or calls it (bench.py's cpu_baseline leg excepted).
 *
 * What it restates (reference file:line):
 *   zo_crc32_update / zo_crc32_combine   src/crc32.h:22-38, 48-65       (IEEE CRC-32, zlib's GF(2) combine)
 *   zo_normalize                          src/codec.h:41-66              (counts -> 12-bit frequencies)
 *   zo_encode_stream / zo_decode_stream   src/codec.h:96-147, 150-205    (64-lane rANS stream section)
 *   zo_cpu_parse                          src/codec2.h:43-81             (hash-chain LZ77, lazy 1)
 *   zo_pack_tokens                        src/codec2.h:84-153            (groups, recent offsets, 5 streams, bits)
 *   zo_decompress_block                   src/codec2.h:155-203           (sequential ZB2 decoder)
 *   zo_gpu_* (the GPU encoder's model)    kernels/zero.cl:49-317         (hash5, sorted buckets, match, lazy,
 *                                                                         greedy chain with LENCAP extension)
 * Parity pins: see the header of each function in zoracle.c.
 */
#ifndef ZORACLE_H
#define ZORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint32_t i, m, d; } zo_token; /* match at position i, length m, distance d */

uint32_t zo_crc32_update(uint32_t crc, const uint8_t* p, size_t n);
uint32_t zo_crc32_combine(uint32_t a, uint32_t b, uint64_t len_b);

void zo_normalize(const uint32_t counts[256], uint32_t total, uint16_t freq[256]);
/* one stream section (u8 kind, u32 len, body) into out; returns its size. out needs n + 5 + 400 bytes. */
size_t zo_encode_stream(const uint8_t* sym, uint32_t n, uint8_t* out);
/* decodes a section of avail bytes holding n symbols; returns the bytes consumed, or a negative error */
long zo_decode_stream(const uint8_t* p, size_t avail, uint32_t n, uint8_t* sym);

/* CPU parse with hash-chain depth `depth`; returns the token count (tokens needs n/4 + 1 entries) */
size_t zo_cpu_parse(const uint8_t* in, uint32_t n, int depth, zo_token* tokens);
/* ZB2 payload from a token list; returns its size (out needs zo_payload_cap(n) bytes) */
size_t zo_pack_tokens(const uint8_t* in, uint32_t n, const zo_token* tk, size_t ntk, uint8_t* out);
size_t zo_payload_cap(uint32_t n);
/* the reference CPU block compressor: payload size; *stored = 1 when the payload is not smaller than n */
size_t zo_cpu_compress_block(const uint8_t* in, uint32_t n, int depth, uint8_t* out, int* stored);
/* 0 on success, a negative code on damage */
int zo_decompress_block(const uint8_t* in, size_t in_len, uint8_t* out, uint32_t n);

/* the GPU encoder's model. level 1..9 -> (NCANDL, LENCAP) = (2,16) L<=2, (4,16) L<=6, (8,32) else */
void zo_gpu_level(int level, uint32_t* ncand, uint32_t* lencap);
/* match[p] = len << 20 | (dist - 1), or 0 (match needs n entries) */
void zo_gpu_match(const uint8_t* in, uint32_t n, uint32_t block_log, int level, uint32_t* match);
/* lazy[p] from match (both n entries) */
void zo_gpu_lazy(const uint32_t* match, uint32_t n, uint32_t* lazy);
/* greedy chain over lazy (extended in place like the GPU); returns the token count */
size_t zo_gpu_parse(const uint8_t* in, uint32_t n, int level, uint32_t* lazy, zo_token* tokens);
size_t zo_gpu_compress_block(const uint8_t* in, uint32_t n, uint32_t block_log, int level, uint8_t* out, int* stored);

#ifdef __cplusplus
}
#endif
#endif
