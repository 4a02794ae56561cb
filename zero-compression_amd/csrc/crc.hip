// Per-block CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320, init and final xor 0xFFFFFFFF), the checksum the
// .zero container stores for every block (src/archive.h:376,618; src/crc32.h:22-39). In the reference the host
// computes it; here one wavefront computes a block's.
//
// Lane i takes the i-th of 64 contiguous chunks (slice-by-8 tables in LDS) starting from a zero register. The raw
// register is linear, so R(block) = S_N(~0) xor (xor over i of S_{after_i}(R(chunk_i))), with S_k(v) = v * x^(8k)
// mod P; x^(8k) comes from a table of x^(8 m 256^j) (m < 256, j < 3) built on the host.
#include "zcommon.h"

namespace zk {

__device__ inline uint32_t gf2_mulmod(uint32_t a, uint32_t b) {  // reflected: bit 31 is x^0
  uint32_t r = 0;
  for (int i = 31; i >= 0; --i) {
    r ^= ((a >> i) & 1u) ? b : 0u;
    b = (b >> 1) ^ ((b & 1u) ? 0xEDB88320u : 0u);
  }
  return r;
}
__device__ inline uint32_t xpow8(const uint32_t* __restrict__ xp, uint32_t k) {  // x^(8k) mod P, k < 2^24
  uint32_t r = xp[k & 255u];
  r = gf2_mulmod(r, xp[256 + ((k >> 8) & 255u)]);
  return gf2_mulmod(r, xp[512 + ((k >> 16) & 255u)]);
}

// crc[b] of the bytes at data + (use_dst ? dst : src), raw_len long, for 4 blocks per workgroup
__global__ __launch_bounds__(256) void k_crc32(const Block* __restrict__ blocks, uint32_t nb, const uint8_t* data,
                                               int use_dst, const uint32_t* __restrict__ xp,
                                               uint32_t* __restrict__ crc) {
  __shared__ uint32_t T[8][256];
  const uint32_t tid = threadIdx.x;
  {
    uint32_t c = tid;
    for (int k = 0; k < 8; ++k) c = (c >> 1) ^ ((c & 1u) ? 0xEDB88320u : 0u);
    T[0][tid] = c;
  }
  __syncthreads();
  {
    uint32_t c = T[0][tid];
    for (int k = 1; k < 8; ++k) { c = (c >> 8) ^ T[0][c & 255u]; T[k][tid] = c; }
  }
  __syncthreads();
  const uint32_t b = blockIdx.x * 4 + (tid >> 6), lane = tid & 63u;
  if (b >= nb) return;
  const Block blk = blocks[b];
  const uint8_t* p = data + (use_dst ? blk.dst : blk.src);
  const uint32_t len = blk.raw_len;
  const uint32_t chunk = (((len + 63u) / 64u) + 7u) & ~7u;
  const uint32_t lo = min(lane * chunk, len), hi = min(lo + chunk, len);
  uint32_t r = 0;
  uint32_t i = lo;
  for (; i + 8 <= hi; i += 8) {
    uint64_t w;
    __builtin_memcpy(&w, p + i, 8);
    const uint32_t a = (uint32_t)w ^ r, c = (uint32_t)(w >> 32);
    r = T[7][a & 255u] ^ T[6][(a >> 8) & 255u] ^ T[5][(a >> 16) & 255u] ^ T[4][a >> 24] ^ T[3][c & 255u] ^
        T[2][(c >> 8) & 255u] ^ T[1][(c >> 16) & 255u] ^ T[0][c >> 24];
  }
  for (; i < hi; ++i) r = T[0][(r ^ p[i]) & 255u] ^ (r >> 8);
  uint32_t v = hi > lo ? gf2_mulmod(r, xpow8(xp, len - hi)) : 0u;
  if (lane == 0) v ^= gf2_mulmod(0xFFFFFFFFu, xpow8(xp, len));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v ^= (uint32_t)__shfl_xor(v, o, 64);
  if (lane == 0) crc[b] = ~v;
}

}  // namespace zk
