// ZB2 block encoder for gfx950. The algorithm is the reference GPU encoder's (kernels/zero.cl:44-618) made
// deterministic: every bucket is sorted completely by (key, position), so the match candidates of a position are
// exactly the nearest earlier positions with its 27-bit hash key, whatever order the scatter's atomics produced.
//
// Per launch: nb blocks, block b's raw bytes at in + blocks[b].src (blocks[b].raw_len of them). Per-block scratch
// is indexed by b (EncLayout in capi.hip). Stages:
//   k_enc_hash      hash5 of every position, bucket counts (16-bit, in LDS) and bstart, the positions grouped by
//                   partition
//   k_enc_place     a partition's positions into their buckets, through LDS
//   k_enc_sort_s    buckets of <= 16 entries sorted in registers, <= 64 across a wavefront; larger ones queued
//   k_enc_sort_m    queued buckets of 65..4096 entries: bitonic sort in LDS
//   k_enc_sort_l    larger buckets: the dominant key's entries through a bitmap of positions in LDS, the few others
//                   ranked among themselves; with many others, a stable LSD radix sort through global memory
//   k_enc_match     the best match of every position among its run check and sorted neighbours
//   k_enc_parse     the greedy chain (with lazy matching): segment walkers in parallel, then a join of their walks
//   k_enc_tokens    the chain -> literals and tokens; groups, recent offsets, tok/ext/sym/groups, offset bits
//   k_enc_rans      per (block, stream): histogram, 12-bit frequencies, 64 lanes of rANS
//   k_enc_pack      the payload, or "stored" when it is not smaller than the block
//
// Each stage's body is a __forceinline__ function of (block, item), which the plain kernels here and the persistent
// ones in plan.hip both call: a body that two kernels share is otherwise compiled as a call, and the call's register
// convention more than doubles a kernel's VGPRs (so halves its occupancy or worse).
#include "zcommon.h"

namespace zk {

#define KEY_TAG_BITS 12u
__host__ __device__ inline uint32_t hbits_of(uint32_t blog) { return blog - 3u < 15u ? blog - 3u : 15u; }
// the parse's window of matches per walker (entries; 32, a whole 128-byte line, was slower on pages with long
// matches, whose walks use few entries of a window)
constexpr uint32_t PARSE_WIN = 16;
// the parse's LDS words for `walkers` segments: per walker its exit, entry, meeting point and a window (PARSE_WIN + 1
// words apart, so that the walkers' window accesses spread over the banks). The chain bitmap is gathered in the
// block's mark words in HBM, each walker storing its own segment's words, so that it (32 KiB at 2^18, 128 KiB at
// 2^20) does not limit the workgroups per CU.
__host__ __device__ inline size_t parse_lds_words(uint32_t blog, uint32_t walkers) {
  (void)blog;
  return (4 + PARSE_WIN) * (size_t)walkers;
}
constexpr uint32_t PARSE_SPAN = 1024;       // default parse segment: positions per walker
constexpr uint32_t PARSE_WALKERS_MIN = 128;  // default walkers per block, at least (each segment still >= 32 positions)

__device__ inline uint32_t ld32(const uint8_t* p) { uint32_t v; __builtin_memcpy(&v, p, 4); return v; }
__device__ inline uint64_t ld64(const uint8_t* p) { uint64_t v; __builtin_memcpy(&v, p, 8); return v; }

// hash of 5 bytes: v the first 4 (little endian), w the fifth, read as 0 past the block (kernels/zero.cl:49-58)
__device__ inline uint32_t hash5v(uint32_t v, uint32_t w) { return (v * 2654435761u) ^ (w * 0x9E3779B1u + (w >> 3)); }

// common prefix of the bytes at a and b (a < b), up to maxlen; b + maxlen <= n keeps every read in the block
__device__ inline uint32_t prefix_len(const uint8_t* inb, uint32_t a, uint32_t b, uint32_t maxlen) {
  uint32_t l = 0;
  while (l + 8 <= maxlen) {
    const uint64_t x = ld64(inb + a + l) ^ ld64(inb + b + l);
    if (x) return l + ((uint32_t)__builtin_ctzll(x) >> 3);
    l += 8;
  }
  if (l + 4 <= maxlen) {
    const uint32_t x = ld32(inb + a + l) ^ ld32(inb + b + l);
    if (x) return l + ((uint32_t)__builtin_ctz(x) >> 3);
    l += 4;
  }
  while (l < maxlen && inb[a + l] == inb[b + l]) ++l;
  return l;
}

__device__ inline uint32_t bitlen(uint32_t d) { return 31u - (uint32_t)__builtin_clz(d); }

// The dispatcher hands workgroup i to XCD i % 8. A position-parallel kernel reads and writes one block's arrays
// (1 MiB each at 2^18) from a thousand workgroups; renumbered so that each XCD takes a contiguous run of them, a
// block's lines are filled in one XCD's L2 instead of partly in all eight.
__device__ inline uint32_t xcd_swizzle(uint32_t i, uint32_t G) {
  const uint32_t per = G / 8u, rem = G % 8u, x = i % 8u, j = i / 8u;
  return x * per + min(x, rem) + j;
}

// level 1..9 -> (NCANDL, LENCAP): neighbours compared and bytes compared before the extension (src/gpu_codec.h:66-70)
__host__ __device__ inline void level_params(uint32_t level, uint32_t& ncand, uint32_t& lencap) {
  ncand = level <= 2 ? 2u : level <= 6 ? 4u : 8u;
  lencap = level <= 6 ? 16u : 32u;
}

// the 8 bytes from q (bytes past the block read as 0), one load at a clamped address; n >= 8
__device__ inline uint64_t ld8z(const uint8_t* inb, uint32_t n, uint32_t q) {
  const uint32_t qq = min(q, n), a = min(qq, n - 8u), sh = 8u * (qq - a);
  const uint64_t x = ld64(inb + a);
  return sh >= 64u ? 0ull : x >> sh;
}
__device__ inline uint64_t ld8z_small(const uint8_t* inb, uint32_t n, uint32_t q) {  // any n
  uint64_t x = 0;
  for (uint32_t t = 0; t < 8; ++t) if (q + t < n) x |= (uint64_t)inb[q + t] << (8 * t);
  return x;
}
// hash5 of position q + j from the 8 bytes x at q (j < 4)
__device__ inline uint32_t hash5_of(uint64_t x, uint32_t j) {
  return hash5v((uint32_t)(x >> (8 * j)), (uint32_t)(x >> (8 * j + 32)) & 0xFFu);
}

// ---------------------------------------------------------------- 1. hash, count, scatter
// The key of a position is the top `keybits` bits of its hash; buckets are its top `hbits` bits. srt[b*B + slot] =
// (the key's other keybits - hbits bits) << blog | position, grouped by bucket; bstart[b*(HB+1) + h] = the first slot
// of bucket h. Once each bucket is sorted, positions are in (key, position) order whatever hbits is, so hbits (and
// the slab size 2^blog, any size that holds the block) changes the work, not the candidates.
//
// Two kernels, so that no write lands at a random place in a block's 4-byte-per-position arrays (such writes cost
// whole lines of HBM traffic each): k_enc_hash (one workgroup per block) counts the buckets, writes bstart, and puts
// the positions in partition order into tmp, a partition being the buckets that share the top pbits hash bits
// (about 8192 positions); k_enc_place (one workgroup per partition) puts a partition's entries in bucket order in LDS
// and writes them out in order.
constexpr uint32_t PLACE_CAP = 15360;  // entries of a partition placed in LDS (larger ones go straight to srt)
__host__ __device__ inline uint32_t pbits_of(uint32_t blog, uint32_t hbits) {
  const uint32_t want = blog > 13u ? blog - 13u : 0u;
  return want < hbits ? want : hbits;
}
// LDS counters that many lanes of a wavefront may hit at once: the positions of a long run of one byte value all hash
// alike, and atomics of many lanes on one address are served one after another. crowded(): whether more than 16 of
// the active lanes share the index of the first of them (a hint, taken once for several atomics); lds_inc(.., true)
// then gives the lanes where `on` that share the first one's index one atomic between them, ranked by lane. Either
// form returns each lane's old value of a[i] (after an atomicAdd of 1).
__device__ inline bool crowded(uint32_t i) {
  return popc64(ballot(i == (uint32_t)__builtin_amdgcn_readfirstlane((int)i))) > 16u;
}
__device__ inline uint32_t lds_inc(uint32_t* a, uint32_t i, bool on, bool crowd) {
  if (!crowd) return on ? atomicAdd(&a[i], 1u) : 0u;  // (uniform)
  const uint64_t act = ballot(on);
  if (!act) return 0u;
  const uint32_t lead = (uint32_t)__builtin_ctzll(act);
  const uint32_t first = readlane(i, lead);
  const bool same = on && i == first;
  const uint64_t sm = ballot(same);
  uint32_t b = 0;
  if (lane_id() == lead) b = atomicAdd(&a[first], popc64(sm));
  b = readlane(b, lead);
  if (same) return b + mbcnt64(sm);
  return on ? atomicAdd(&a[i], 1u) : 0u;
}
// k_enc_hash's bucket counters, two 16-bit counters to a word (64 KiB for 2^15 buckets: two workgroups per CU). A
// count past 65535 is kept exact by a log of the atomics that carried: one whose low half passed a multiple of 65536
// (the word's index) and one whose word passed 2^32 (the index | 1 << 31). With adds of at most 64 each carry is one
// crossing, and a block of 2^20 positions makes at most 16 of each kind. The true counts of word w: low =
// (w & 0xFFFF) + 65536 x (its low carries), high = (2^32 x (its word carries) + w - low) >> 16.
constexpr uint32_t HASH_CARRY_MAX = 32;
__device__ inline void lds_count16(uint32_t* a, uint32_t* carries, uint32_t i, bool on, bool crowd) {
  uint32_t w = i >> 1, add = 1u << ((i & 1u) << 4), old = 0;
  bool did = false;
  if (crowd) {  // (uniform) the lanes on the first lane's counter as one atomic, the others one each
    const uint64_t act = ballot(on);
    if (!act) return;
    const uint32_t lead = (uint32_t)__builtin_ctzll(act);
    const uint32_t first = readlane(i, lead);
    const bool same = on && i == first;
    const uint32_t k = popc64(ballot(same));
    if (lane_id() == lead) {
      add = k << ((first & 1u) << 4);
      old = atomicAdd(&a[w], add);
      did = true;
    } else if (on && !same) {
      old = atomicAdd(&a[w], add);
      did = true;
    }
  } else if (on) {
    old = atomicAdd(&a[w], add);
    did = true;
  }
  if (did) {
    if ((old & 0xFFFFu) + (add & 0xFFFFu) > 0xFFFFu) {
      const uint32_t e = atomicAdd(&carries[0], 1u);
      if (e < HASH_CARRY_MAX) carries[1 + e] = w;
    }
    if ((uint64_t)old + add > 0xFFFFFFFFull) {
      const uint32_t e = atomicAdd(&carries[0], 1u);
      if (e < HASH_CARRY_MAX) carries[1 + e] = w | 0x80000000u;
    }
  }
}
// the true count of counter i (0 where !in), the wavefront's lanes taking the 64 consecutive counters from i - lane
// (all lanes active); ev: lane e's entry of the carry log, ~0 past its end. Only the entries on the step's words are
// looked at.
__device__ inline uint32_t count16_at(const uint32_t* a, uint32_t ev, uint32_t i, bool in) {
  const uint32_t w = in ? a[i >> 1] : 0u;
  const uint32_t base = (i - lane_id()) >> 1;
  uint64_t hit = ballot(((ev & 0x7FFFFFFFu) - base) <= 32u);  // (words base .. base + 32: i - lane may be odd)
  uint32_t lc = 0, wc = 0;
  for (; hit; hit &= hit - 1) {
    const uint32_t v = readlane(ev, (uint32_t)__builtin_ctzll(hit));
    if ((v & 0x7FFFFFFFu) == (i >> 1)) {
      if (v >> 31) ++wc;
      else ++lc;
    }
  }
  if (!in) return 0u;
  const uint32_t lo = (w & 0xFFFFu) + (lc << 16);
  return (i & 1u) ? (uint32_t)((((uint64_t)wc << 32) + w - lo) >> 16) : lo;
}
// k_enc_hash's LDS: the bucket counters (whose words, once scanned, hold the 16 wavefronts' staging and then their
// partition counts and bases), 16 wavefront sums, a flag, the partition cursors and the carry log
constexpr uint32_t HASH_STAGE = 256;  // positions a wavefront stages per step
__host__ __device__ inline uint32_t hash_counter_words(uint32_t hbits, uint32_t pbits) {
  const uint32_t later = 16u * HASH_STAGE + 32u * (1u << pbits);
  return (1u << hbits) / 2u > later ? (1u << hbits) / 2u : later;
}
__host__ __device__ inline size_t hash_lds_words(uint32_t hbits, uint32_t pbits) {
  return (size_t)hash_counter_words(hbits, pbits) + 17 + ((size_t)1 << pbits) + 1 + HASH_CARRY_MAX;
}
__device__ __forceinline__ void enc_hash_block(const Block* __restrict__ blocks, const uint8_t* __restrict__ in, uint32_t blog,
                                      uint32_t hbits, uint32_t pbits, uint32_t* __restrict__ tmp,
                                      uint32_t* __restrict__ bstart, uint32_t b) {
  extern __shared__ uint32_t cnt[];  // hash_lds_words(hbits, pbits)
  const uint32_t tid = threadIdx.x;
  const uint32_t HBITS = hbits, HB = 1u << HBITS, CW = hash_counter_words(hbits, pbits);
  const Block blk = blocks[b];
  const uint8_t* inb = in + blk.src;
  const uint32_t n = enc_len(blk, blog);
  const uint32_t count = n >= 4 ? n - 3 : 0;
  uint32_t* bigb = cnt + CW + 16;  // (after the 16 wavefront sums) a bucket of more than 4096 entries exists
  const uint32_t NP = 1u << pbits, pmask = (1u << blog) - 1u;
  uint32_t* pcur = cnt + CW + 17;                   // [NP] next free slot of each partition
  uint32_t* carries = cnt + CW + 17 + NP;           // [1 + HASH_CARRY_MAX] the counters' carry log
  for (uint32_t i = tid; i < HB / 2u; i += 1024) cnt[i] = 0;
  if (tid == 0) { *bigb = 0u; carries[0] = 0u; }
  __syncthreads();
  // 4 positions per lane from one 8-byte load, the next load in flight while these are counted
  {
    uint64_t x = n >= 8 ? ld8z(inb, n, 4 * tid) : ld8z_small(inb, n, 4 * tid);
    for (uint32_t q = 4 * tid; q < count; q += 4096) {
      const uint64_t nx = n >= 8 ? ld8z(inb, n, q + 4096) : 0ull;
      const bool crowd = crowded(hash5_of(x, 0) >> (32 - HBITS));
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) lds_count16(cnt, carries, hash5_of(x, j) >> (32 - HBITS), q + j < count, crowd);
      x = nx;
    }
  }
  __syncthreads();
  // exclusive scan of the HB counters, HB/16 per wavefront, 64 consecutive counters per step; the bucket starts go to
  // bstart, the partitions' first ones to their cursors
  const uint32_t wv = tid >> 6, lane = tid & 63u, chunk = HB / 16u;
  const uint32_t ev = lane < min(carries[0], HASH_CARRY_MAX) ? carries[1 + lane] : ~0u;
  uint32_t* wsum = cnt + CW;  // [16]
  {
    uint32_t s = 0;
    for (uint32_t i0 = 0; i0 < chunk; i0 += 64) s += count16_at(cnt, ev, wv * chunk + i0 + lane, i0 + lane < chunk);
    s = wave_sum(s);
    if (lane == 0) wsum[wv] = s;
  }
  __syncthreads();
  uint32_t* bs = bstart + (size_t)b * (HB + 1);
  {
    uint32_t run = 0;
    for (uint32_t x = 0; x < wv; ++x) run += wsum[x];
    for (uint32_t i0 = 0; i0 < chunk; i0 += 64) {
      const uint32_t i = wv * chunk + i0 + lane;
      const bool in = i0 + lane < chunk;
      const uint32_t c = count16_at(cnt, ev, i, in);
      const uint32_t inc = wave_scan_incl(c);
      if (in) {
        bs[i] = run + inc - c;
        if ((i & ((1u << (HBITS - pbits)) - 1u)) == 0) pcur[i >> (HBITS - pbits)] = run + inc - c;
      }
      if (in && c > 4096u) *bigb = 1u;  // (lanes that write write the same value)
      run += __builtin_amdgcn_readlane(inc, 63);
    }
  }
  if (tid == 0) bs[HB] = count;
  __syncthreads();
  const bool big = *bigb != 0u;  // a bucket of more than 4096: a long run, whose positions crowd the counters
  // the positions grouped by partition (in no set order inside one: k_enc_place does not need it). A wavefront takes
  // 256 positions a step: LDS counters of its own rank them per partition (the atomics' return values), one claim
  // per partition from the shared cursors places each partition's run, and the positions, put in partition order in
  // the wavefront's staging, go out as runs of consecutive slots (a store instruction covers a few runs, instead of
  // about one slot in every partition: each slot written at a scattered place costs an L2 request of its own). No
  // workgroup barrier inside the loop.
  uint32_t* wc = cnt + 16u * HASH_STAGE + wv * 2u * NP;  // [16][2 NP] (counter words): per partition this step's
                                                         // count, then its offset among the step's positions;
                                                         // [NP + i]: its first slot minus that offset
  uint32_t* wst = cnt + wv * HASH_STAGE;     // [16][HASH_STAGE] (the counters' words, no longer needed)
  uint32_t* tb = tmp + ((size_t)b << blog);
  uint64_t x = n >= 8 ? ld8z(inb, n, 4 * tid) : ld8z_small(inb, n, 4 * tid);
  for (uint32_t q = 4 * tid; q - 4 * lane < count; q += 4096) {  // (uniform: a wavefront runs while any lane has work)
    const uint64_t nx = n >= 8 ? ld8z(inb, n, q + 4096) : 0ull;
    for (uint32_t i = lane; i < NP; i += 64) wc[i] = 0;
    wave_fence();
    uint32_t d[4], rank[4];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) d[j] = pbits ? hash5_of(x, j) >> (32 - pbits) : 0u;
    const bool crowd = big && crowded(d[0]);
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) rank[j] = lds_inc(wc, d[j], q + j < count, crowd);
    wave_fence();
    uint32_t run = 0;  // (after the loop: the step's positions)
    for (uint32_t i0 = 0; i0 < NP; i0 += 64) {
      const uint32_t i = i0 + lane;
      const uint32_t c = i < NP ? wc[i] : 0u;
      const uint32_t inc = wave_scan_incl(c);
      if (i < NP) {
        const uint32_t off = run + inc - c;
        const uint32_t first = c ? atomicAdd(&pcur[i], c) : 0u;
        wc[i] = off;
        wc[NP + i] = first - off;
      }
      run += readlane(inc, 63);
    }
    wave_fence();
    if (pbits >= 4) {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j)
        if (q + j < count) wst[wc[d[j]] + rank[j]] = d[j] << blog | (q + j);
      wave_fence();
      for (uint32_t L = lane; L < run; L += 64) {
        const uint32_t v = wst[L];
        tb[wc[NP + (v >> blog)] + L] = v & pmask;
      }
    } else {  // (fewer than 16 partitions: the runs are long enough as the lanes hold them)
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j)
        if (q + j < count) tb[wc[NP + d[j]] + wc[d[j]] + rank[j]] = q + j;
    }
    wave_fence();  // (the next step's counts overwrite wc, its positions wst)
    x = nx;
  }
}
__global__ __launch_bounds__(1024) void k_enc_hash(const Block* __restrict__ blocks, const uint8_t* __restrict__ in,
                                                   uint32_t blog, uint32_t hbits, uint32_t pbits,
                                                   uint32_t* __restrict__ tmp, uint32_t* __restrict__ bstart) {
  enc_hash_block(blocks, in, blog, hbits, pbits, tmp, bstart, blockIdx.x);
}

// one workgroup per (block, partition): the partition's positions (from tmp) to their buckets
__device__ __forceinline__ void enc_place_part(const Block* __restrict__ blocks, const uint8_t* __restrict__ in, uint32_t blog,
                                      uint32_t hbits, uint32_t keybits, uint32_t pbits,
                                      const uint32_t* __restrict__ tmp, const uint32_t* __restrict__ bstart,
                                      uint32_t* __restrict__ srt, uint32_t b, uint32_t P) {
  extern __shared__ uint32_t sm[];  // bucket cursors (2^(hbits - pbits)), then the partition's entries
  const uint32_t tid = threadIdx.x;
  const uint32_t HB = 1u << hbits, pb = hbits - pbits, NB = 1u << pb;
  const Block blk = blocks[b];
  const uint8_t* inb = in + blk.src;
  const uint32_t n = enc_len(blk, blog);
  const uint32_t* bs = bstart + (size_t)b * (HB + 1);
  const uint32_t ps = bs[P << pb], pe = bs[(P + 1) << pb], m = pe - ps;
  uint32_t* lc = sm;
  uint32_t* buf = sm + NB;
  for (uint32_t i = tid; i < NB; i += 1024) lc[i] = bs[(P << pb) + i] - ps;
  __syncthreads();
  const bool in_lds = m <= PLACE_CAP;
  const uint32_t* tb = tmp + ((size_t)b << blog) + ps;
  uint32_t* sb = srt + ((size_t)b << blog) + ps;
  const uint32_t tagmask = (1u << (keybits - hbits)) - 1u;
  for (uint32_t i0 = tid; i0 < m; i0 += 4096) {  // 4 entries per lane, every load issued before any is used
    uint32_t p[4];
    uint64_t x[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) p[u] = tb[min(i0 + 1024 * u, m - 1)];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) x[u] = n >= 8 ? ld8z(inb, n, p[u]) : ld8z_small(inb, n, p[u]);
    // (a partition too large for LDS holds a long run's bucket)
    const bool crowd = !in_lds && crowded((hash5_of(x[0], 0) >> (32 - hbits)) & (NB - 1u));
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u)
      if (i0 + 1024 * u < m) {
        const uint32_t h = hash5_of(x[u], 0);
        const uint32_t slot = lds_inc(lc, (h >> (32 - hbits)) & (NB - 1u), true, crowd);
        const uint32_t e = ((h >> (32 - keybits)) & tagmask) << blog | p[u];
        if (in_lds) buf[slot] = e;
        else sb[slot] = e;
      }
  }
  __syncthreads();
  if (in_lds)
    for (uint32_t i = tid; i < m; i += 1024) sb[i] = buf[i];
}
__global__ __launch_bounds__(1024) void k_enc_place(const Block* __restrict__ blocks, const uint8_t* __restrict__ in,
                                                    uint32_t nb, uint32_t blog, uint32_t hbits, uint32_t keybits,
                                                    uint32_t pbits, const uint32_t* __restrict__ tmp,
                                                    const uint32_t* __restrict__ bstart, uint32_t* __restrict__ srt) {
  const uint32_t wg = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint32_t b = wg >> pbits, P = wg & ((1u << pbits) - 1u);
  if (b >= nb) return;  // (uniform over the workgroup)
  enc_place_part(blocks, in, blog, hbits, keybits, pbits, tmp, bstart, srt, b, P);
}

// ---------------------------------------------------------------- 2. bucket sorts (the values are distinct)
struct SortQueues {
  uint32_t* mid;    // (block, bucket) pairs of 65..4096 entries
  uint32_t* large;  // pairs of more than 4096 entries
  uint32_t* counts; // [0] mid, [1] large
  uint32_t cap;     // pairs each queue can hold
};

// Buckets of <= 16 entries: one thread each, a sorting network in registers. Buckets of 17..64: the thread's
// wavefront sorts them one after another, one entry per lane, a bitonic network over lane shuffles. Larger buckets
// are queued for k_enc_sort_m (<= 4096) or k_enc_sort_l.
// the k <= 16 values at s sorted (one thread): Batcher's odd-even merge sort of 16, 63 comparators, each a min and a
// max (the padding sorts last). The lanes of a wavefront sort 64 buckets at scattered places, so each access costs
// an L2 request whatever its width up to 16 bytes: the bucket is read 16 bytes at a time (the last read may take up
// to 3 entries past it, which are not written), and written back 16 bytes at a time where 4 entries remain.
__device__ inline void sort16(uint32_t* s, uint32_t k) {
  uint32_t v[16];
#pragma unroll
  for (int q = 0; q < 16; q += 4) {
    uint32_t t[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    if ((uint32_t)q < k) __builtin_memcpy(t, s + q, 16);
    v[q] = t[0]; v[q + 1] = t[1]; v[q + 2] = t[2]; v[q + 3] = t[3];
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = (uint32_t)i < k ? v[i] : 0xFFFFFFFFu;
#pragma unroll
  for (int p = 1; p < 16; p <<= 1)
#pragma unroll
    for (int kk = p; kk >= 1; kk >>= 1)
#pragma unroll
      for (int j = kk % p; j + kk < 16; j += 2 * kk)
#pragma unroll
        for (int i = 0; i < kk && i + j + kk < 16; ++i)
          if ((i + j) / (2 * p) == (i + j + kk) / (2 * p)) {
            const uint32_t a = v[i + j], c = v[i + j + kk];
            v[i + j] = min(a, c);
            v[i + j + kk] = max(a, c);
          }
#pragma unroll
  for (int q = 0; q < 16; q += 4) {
    if ((uint32_t)q + 4u <= k) {
      const uint32_t t[4] = {v[q], v[q + 1], v[q + 2], v[q + 3]};
      __builtin_memcpy(s + q, t, 16);
    } else if ((uint32_t)q < k) {
      s[q] = v[q];
      if ((uint32_t)q + 1u < k) s[q + 1] = v[q + 1];
      if ((uint32_t)q + 2u < k) s[q + 2] = v[q + 2];
    }
  }
}
// the value of lane ^ j (j = 1, 2, ..., 32) without LDS traffic: DPP inside 16-lane rows (quad permutations, row
// rotations: row_ror:n gives lane i the value of lane i - n of its row), the gfx950 permlane swaps across rows
// (permlane16_swap(v, v) gives [v of lane i & ~16, v of lane i | 16], permlane32_swap likewise for bit 5)
__device__ inline uint32_t xor_lane(uint32_t v, uint32_t j, uint32_t lane) {
  switch (j) {
    case 1: return dpp0<0xB1>(v);  // quad_perm [1,0,3,2]
    case 2: return dpp0<0x4E>(v);  // quad_perm [2,3,0,1]
    case 4: { const uint32_t up = dpp0<0x12C>(v), dn = dpp0<0x124>(v); return (lane & 4u) ? dn : up; }
    case 8: return dpp0<0x128>(v);  // row_ror:8
    case 16: { const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false); return (lane & 16u) ? r[0] : r[1]; }
    default: { const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false); return (lane & 32u) ? r[0] : r[1]; }
  }
}
// the k (17..64) values at s sorted by the whole wavefront, one per lane: a bitonic network over lane exchanges
__device__ inline void sort64_wave(uint32_t* s, uint32_t k) {
  const uint32_t lane = lane_id();
  uint32_t v = lane < k ? s[lane] : 0xFFFFFFFFu;
#pragma unroll
  for (uint32_t kk = 2; kk <= 64; kk <<= 1)
#pragma unroll
    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
      const uint32_t o = xor_lane(v, j, lane);
      const bool keep_min = ((lane & j) == 0) == ((lane & kk) == 0);
      v = keep_min ? min(v, o) : max(v, o);
    }
  if (lane < k) s[lane] = v;
}
// two buckets of at most 32 values sorted side by side, the k values at s by lanes 0..31 and the k2 at s2 by lanes
// 32..63 (k2 = 0: none): the network of 32 (15 steps instead of 21), whose exchanges stay inside each half, every
// half sorting upwards
__device__ inline void sort32x2_wave(uint32_t* s, uint32_t k, uint32_t* s2, uint32_t k2) {
  const uint32_t lane = lane_id(), l = lane & 31u;
  uint32_t* p = lane < 32u ? s : s2;
  const uint32_t kk2 = lane < 32u ? k : k2;
  uint32_t v = l < kk2 ? p[l] : 0xFFFFFFFFu;
#pragma unroll
  for (uint32_t kk = 2; kk <= 32; kk <<= 1)
#pragma unroll
    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
      const uint32_t o = xor_lane(v, j, lane);
      const bool keep_min = ((l & j) == 0) == ((l & kk) == 0);
      v = keep_min ? min(v, o) : max(v, o);
    }
  if (l < kk2) p[l] = v;
}
// bucket h of block b queued for the larger sorts when it has more than 64 entries
__device__ inline void queue_large(uint32_t k, uint32_t b, uint32_t h, SortQueues q) {
  if (k > 64) {
    const uint32_t which = k > 4096 ? 1 : 0;
    const uint32_t i = atomicAdd(&q.counts[which], 1u);
    if (i < q.cap) { uint32_t* dst = which ? q.large : q.mid; dst[2 * i] = b; dst[2 * i + 1] = h; }
  }
}
// the bucket [lo, lo + k) of each thread (k <= 64; larger ones are left as they are), at base + (lo - at)
__device__ inline void sort_small(uint32_t* base, uint32_t at, uint32_t lo, uint32_t k) {
  if (k >= 2 && k <= 16) sort16(base + (lo - at), k);
  uint64_t m = ballot(k > 16 && k <= 32);  // two at a time
  while (m) {
    const uint32_t a = (uint32_t)__builtin_ctzll(m);
    m &= m - 1;
    const uint32_t c = m ? (uint32_t)__builtin_ctzll(m) : a;
    m &= m - 1;  // (0 stays 0)
    sort32x2_wave(base + (readlane(lo, a) - at), readlane(k, a), base + (readlane(lo, c) - at),
                  c != a ? readlane(k, c) : 0u);
  }
  m = ballot(k > 32 && k <= 64);
  while (m) {
    const uint32_t who = (uint32_t)__builtin_ctzll(m);
    m &= m - 1;
    sort64_wave(base + (readlane(lo, who) - at), readlane(k, who));
  }
}

// bucket h of block b on this thread (valid false: a thread with no bucket, which still takes part in the
// wavefront's sorts of the others). A wavefront's buckets are consecutive, so their entries are one span of srt: it
// is read into the wavefront's LDS with coalesced loads, the buckets are sorted there, and it is written back with
// coalesced stores (sorted in place, the lanes' scattered bucket accesses each cost an L2 request). A span of more
// than SORT_S_SPAN entries (it holds a bucket queued for the larger sorts), or lanes of more than one block, sort in
// place. Workgroups of 256 threads.
constexpr uint32_t SORT_S_SPAN = 1024;
__device__ __forceinline__ void enc_sort_bucket(uint32_t blog, uint32_t hbits, uint32_t* __restrict__ srt,
                                       const uint32_t* __restrict__ bstart, SortQueues q, uint32_t b, uint32_t h,
                                       bool valid) {
  __shared__ uint32_t span_lds[4][SORT_S_SPAN + 4];  // (+ 4: sort16 reads up to 3 entries past a bucket)
  const uint32_t HB = 1u << hbits, lane = lane_id();
  uint32_t lo = 0, k = 0;
  if (valid) {
    const uint32_t* bs = bstart + (size_t)b * (HB + 1);
    lo = bs[h];
    k = bs[h + 1] - lo;
  }
  if (valid) queue_large(k, b, h, q);
  const uint64_t vm = ballot(valid);
  if (!vm) return;  // (uniform)
  const uint32_t first = (uint32_t)__builtin_ctzll(vm), last = 63u - (uint32_t)__builtin_clzll(vm);
  const uint32_t bf = readlane(b, first);
  uint32_t* sb = srt + ((size_t)bf << blog);
  const uint32_t s0 = readlane(lo, first), span = readlane(lo + k, last) - s0;
  if (ballot(valid && b != bf) || span > SORT_S_SPAN) {  // (uniform)
    sort_small(srt + ((size_t)b << blog), 0, lo, k);
    return;
  }
  uint32_t* lds = span_lds[threadIdx.x >> 6];
  for (uint32_t i = lane; i < span; i += 64) lds[i] = sb[s0 + i];
  wave_fence();
  sort_small(lds, s0, lo, k);
  wave_fence();
  for (uint32_t i = lane; i < span; i += 64) sb[s0 + i] = lds[i];
}
__global__ __launch_bounds__(256) void k_enc_sort_s(const Block* __restrict__ blocks, uint32_t nb, uint32_t blog,
                                                    uint32_t hbits, uint32_t* __restrict__ srt,
                                                    const uint32_t* __restrict__ bstart, SortQueues q) {
  const uint32_t HB = 1u << hbits;
  const uint64_t gid = (uint64_t)xcd_swizzle(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const uint32_t b = (uint32_t)(gid / HB), h = (uint32_t)(gid % HB);
  enc_sort_bucket(blog, hbits, srt, bstart, q, b, h, b < nb);
}

// queued buckets of 65..4096 entries: one workgroup per bucket, bitonic sort in LDS (padded to a power of two)
// the queued buckets w0, w0 + step, ... of q.mid (a workgroup's share of them)
__device__ __forceinline__ void enc_sort_m_range(uint32_t blog, uint32_t hbits, uint32_t* __restrict__ srt,
                                        const uint32_t* __restrict__ bstart, SortQueues q, uint32_t w0, uint32_t step) {
  __shared__ uint32_t a[4096];
  const uint32_t HB = 1u << hbits;
  const uint32_t total = min(q.counts[0], q.cap);
  for (uint32_t w = w0; w < total; w += step) {
    const uint32_t b = q.mid[2 * w], h = q.mid[2 * w + 1];
    const uint32_t* bs = bstart + (size_t)b * (HB + 1);
    const uint32_t lo = bs[h], k = bs[h + 1] - lo;
    uint32_t* s = srt + ((size_t)b << blog) + lo;
    uint32_t np2 = 128;
    while (np2 < k) np2 <<= 1;
    for (uint32_t i = threadIdx.x; i < np2; i += 256) a[i] = i < k ? s[i] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t kk = 2; kk <= np2; kk <<= 1)
      for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < np2; i += 256) {
          const uint32_t l = i ^ j;
          if (l > i) {
            const bool asc = (i & kk) == 0;
            const uint32_t x = a[i], y = a[l];
            if ((x > y) == asc) { a[i] = y; a[l] = x; }
          }
        }
        __syncthreads();
      }
    for (uint32_t i = threadIdx.x; i < k; i += 256) s[i] = a[i];
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void k_enc_sort_m(uint32_t blog, uint32_t hbits, uint32_t* __restrict__ srt,
                                                    const uint32_t* __restrict__ bstart, SortQueues q) {
  enc_sort_m_range(blog, hbits, srt, bstart, q, blockIdx.x, gridDim.x);
}

// queued buckets of more than 4096 entries: one workgroup of 1024 per bucket.
// Such a bucket is mostly one key: a long run of a byte value, or of any short pattern, gives every position in it the
// same 5 bytes, and the bucket holds a few other keys besides (collisions of the bucket's hash bits). The entries of
// the dominant tag are in order once their positions are: they are set in a bitmap of the block's positions (dynamic
// LDS, sort_l_lds_bytes) and read back 64 at a time, each set bit's lane writing its entry at its rank; the others
// (at most SORT_L_OTHERS) are ranked among themselves and go before or after that run. Buckets with more others go
// through a stable LSD radix sort with 8-bit digits through a temporary of the same size (tmp, indexed like srt),
// in-wavefront ranks from ballots.
__host__ __device__ inline size_t sort_l_lds_bytes(uint32_t blog) { return ((size_t)1 << blog) / 8; }
constexpr uint32_t SORT_L_OTHERS = 1024;  // entries of other tags a bucket may hold and still take the bitmap
// the queued buckets w0, w0 + step, ... of q.large
__device__ __forceinline__ void enc_sort_l_range(uint32_t blog, uint32_t hbits, uint32_t* __restrict__ srt,
                                        uint32_t* __restrict__ tmp, const uint32_t* __restrict__ bstart, SortQueues q,
                                        uint32_t w0, uint32_t step) {
  __shared__ uint32_t base[256], tile_tot[256];
  __shared__ uint32_t wcnt[16][256];
  extern __shared__ uint32_t bmp[];  // [2^blog / 32]
  const uint32_t HB = 1u << hbits, B = 1u << blog;
  const uint32_t tid = threadIdx.x, wv = tid >> 6, lane = tid & 63u;
  const uint32_t total = min(q.counts[1], q.cap);
  for (uint32_t w = w0; w < total; w += step) {
    const uint32_t b = q.large[2 * w], h = q.large[2 * w + 1];
    const uint32_t* bs = bstart + (size_t)b * (HB + 1);
    const uint32_t lo = bs[h], k = bs[h + 1] - lo;
    uint32_t* A = srt + ((size_t)b << blog) + lo;
    uint32_t* T = tmp + ((size_t)b << blog) + lo;
    // The tag of the middle entry, td: in a bucket of one dominant key, that key's. Its entries go to the bitmap, the
    // others (collisions of the bucket's hash bits, at most SORT_L_OTHERS of them) to LDS (wcnt's words).
    const uint32_t td = A[k / 2u] >> blog, tdb = td << blog;
    uint32_t* oth = &wcnt[0][0];
    for (uint32_t i = tid; i < B / 32u; i += 1024) bmp[i] = 0;
    if (tid == 0) { base[0] = 0; base[1] = 0; }
    __syncthreads();
    for (uint32_t i0 = 0; i0 < k; i0 += 1024) {  // (uniform)
      const uint32_t i = i0 + tid;
      const uint32_t e = i < k ? A[i] : 0u;
      const bool dom = i < k && (e >> blog) == td, other = i < k && !dom;
      if (dom) atomicOr(&bmp[(e & (B - 1u)) >> 5], 1u << (e & 31u));
      const uint64_t om = ballot(other);
      if (om) {  // (uniform)
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(&base[0], popc64(om));
        at = readlane(at, 0) + mbcnt64(om);
        if (other && at < SORT_L_OTHERS) oth[at] = e;
      }
    }
    __syncthreads();
    const uint32_t m = base[0];
    if (m <= SORT_L_OTHERS) {  // (uniform)
      // the others in order: each one's rank among them, from comparisons with all of them (they are distinct)
      uint32_t ov = 0, orank = 0;
      if (tid < m) {
        ov = oth[tid];
        for (uint32_t j = 0; j < m; ++j) orank += oth[j] < ov ? 1u : 0u;
        if (ov < tdb) atomicAdd(&base[1], 1u);
      }
      // each wavefront's share of the 64-position chunks: its count, then the rank of its first entry
      const uint32_t per = B / 64u / 16u, c0 = wv * per;
      uint32_t s = 0;
      for (uint32_t c = c0 + lane; c < c0 + per; c += 64) s += __popc(bmp[2 * c]) + __popc(bmp[2 * c + 1]);
      s = wave_sum(s);
      if (lane == 0) tile_tot[wv] = s;
      __syncthreads();
      const uint32_t nlt = base[1];  // others before td's entries
      if (tid < m) A[orank < nlt ? orank : orank + (k - m)] = ov;
      uint32_t run = nlt;
      for (uint32_t x = 0; x < wv; ++x) run += tile_tot[x];
      for (uint32_t c = c0; c < c0 + per; ++c) {
        const uint64_t bm = (uint64_t)bmp[2 * c] | (uint64_t)bmp[2 * c + 1] << 32;
        if (bm) {  // (uniform)
          if ((bm >> lane) & 1u) A[run + mbcnt64(bm)] = tdb | (64u * c + lane);
          run += popc64(bm);
        }
      }
      __syncthreads();  // (the next bucket clears the bitmap and reuses base)
      continue;
    }
    for (uint32_t i = tid; i < 16 * 256; i += 1024) (&wcnt[0][0])[i] = 0;
    for (uint32_t pass = 0; pass < 4; ++pass) {
      const uint32_t sh = pass * 8;
      uint32_t* src = (pass & 1) ? T : A;
      uint32_t* dst = (pass & 1) ? A : T;
      if (tid < 256) base[tid] = 0;
      __syncthreads();
      for (uint32_t i = tid; i < k; i += 1024) atomicAdd(&base[(src[i] >> sh) & 255u], 1u);
      __syncthreads();
      if (tid < 64) {  // exclusive scan of 256 counts, 4 per lane
        uint32_t c4[4], s = 0;
        for (int j = 0; j < 4; ++j) { c4[j] = base[tid * 4 + j]; s += c4[j]; }
        uint32_t run = wave_scan_excl(s);
        for (int j = 0; j < 4; ++j) { base[tid * 4 + j] = run; run += c4[j]; }
      }
      __syncthreads();
      for (uint32_t t0 = 0; t0 < k; t0 += 1024) {
        const uint32_t i = t0 + tid;
        const bool valid = i < k;
        const uint32_t v = valid ? src[i] : 0u;
        const uint32_t d = (v >> sh) & 255u;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
          const uint64_t bb = ballot((d >> bit) & 1u);
          m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = popc64(m & lanemask_lt());
        if (valid && rank == 0) wcnt[wv][d] = popc64(m);
        __syncthreads();
        if (tid < 256) {
          uint32_t run = 0;
          for (int x = 0; x < 16; ++x) { const uint32_t c = wcnt[x][tid]; wcnt[x][tid] = run; run += c; }
          tile_tot[tid] = run;
        }
        __syncthreads();
        if (valid) dst[base[d] + wcnt[wv][d] + rank] = v;
        __syncthreads();
        if (tid < 256) {
          base[tid] += tile_tot[tid];
          for (int x = 0; x < 16; ++x) wcnt[x][tid] = 0;
        }
        __syncthreads();
      }
    }
  }
}
__global__ __launch_bounds__(1024) void k_enc_sort_l(uint32_t blog, uint32_t hbits, uint32_t* __restrict__ srt,
                                                     uint32_t* __restrict__ tmp, const uint32_t* __restrict__ bstart,
                                                     SortQueues q) {
  enc_sort_l_range(blog, hbits, srt, tmp, bstart, q, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- 3. match finder (one thread per sorted slot)
// The candidates of a position are the ncand slots just before its own in the sorted order that hold its 12-bit tag
// (the key's low 12 bits) and an earlier position (kernels/zero.cl:186-199); a slot of another bucket with the same
// tag counts too, as there. Only when one exists, the run check (distance 1) goes first (kernels/zero.cl:209-222),
// then each candidate in slot order; the longest wins, the first found on a tie (zero.cl's early reject of a
// candidate whose byte at the best length differs makes its nearer-on-a-tie rule unreachable). A wavefront takes 64
// consecutive slots of one block. The candidates follow from the sorted entries alone (tag << blog | position, the
// tag being the key's bits below the bucket's); then each lane that has a candidate or is one reads the first W bytes
// of its position once (W = 16 for a length cap of 16, 32 for one of 32; bytes past the block read as 0), lanes 0..7
// also those of the 8 slots before the wavefront's first where they are candidates; windows go to LDS, and each
// candidate is compared there, window against window.
constexpr uint32_t MX = 8;             // most candidates of any level
constexpr uint32_t NOPOS = 0xFFFFFFFFu;  // an LDS entry with no slot behind it
constexpr uint32_t MATCH_LDS_WORDS = (MX + 64) * 10;  // a wavefront's entries, flags and windows (NW <= 8)

template <int NW>
__device__ inline void load_window(const uint8_t* inb, uint32_t n, uint32_t p, uint32_t (&w)[NW]) {
  if (p + 4u * NW <= n) {
#pragma unroll
    for (int q = 0; q < NW; q += 4) {
      uint32_t v[4];
      __builtin_memcpy(v, inb + p + 4 * q, 16);
      w[q] = v[0]; w[q + 1] = v[1]; w[q + 2] = v[2]; w[q + 3] = v[3];
    }
  } else {  // the last positions of a block
#pragma unroll
    for (int q = 0; q < NW; ++q) w[q] = 0;
#pragma unroll
    for (int t = 0; t < 4 * NW; ++t)
      if (p + t < n) w[t >> 2] |= (uint32_t)inb[p + t] << (8 * (t & 3));
  }
}


// the number of leading bytes where a and b agree (4 * NW when all do): the first differing bit, 64 bits at a time
// from the last piece (a 64-bit count of trailing zeros that gives 64 for none is two v_ffbl_b32 and a clamped
// add); a piece that agrees whole (64) adds the answer of the pieces after it
template <int NW>
__device__ inline uint32_t common_prefix(const uint32_t (&a)[NW], const uint32_t (&b)[NW]) {
  static_assert(NW % 2 == 0, "64-bit pieces");
  uint32_t lb = 0;
#pragma unroll
  for (int q = NW - 2; q >= 0; q -= 2) {
    const uint64_t x = (uint64_t)(a[q + 1] ^ b[q + 1]) << 32 | (a[q] ^ b[q]);
    const uint32_t t = (uint32_t)__builtin_ctzg(x, 64);
    lb = t + (t >> 6) * lb;
  }
  return lb >> 3;
}

template <int NW>
__device__ inline void match_slots(const uint8_t* __restrict__ inb, uint32_t n, uint32_t count, uint32_t blog,
                                   uint32_t ncand, uint32_t lencap, const uint32_t* __restrict__ sb, uint32_t i0,
                                   uint32_t* __restrict__ mb, uint32_t* sh) {
  const uint32_t lane = threadIdx.x & 63u, i = i0 + lane;
  const uint32_t B = 1u << blog;
  uint32_t* lent = sh;                  // [MX + 64]: entry MX + l is lane l's slot (tag << blog | position), entry
                                        // x < MX slot i0 - MX + x; NOPOS for none
  uint32_t* lneed = sh + MX + 64;       // [MX + 64]: the slot's window is some lane's candidate
  uint32_t* lwin = sh + 2 * (MX + 64);  // [MX + 64][NW]
  const bool xact = lane < MX && i0 + lane >= MX;
  const uint32_t e = i < count ? sb[i] : NOPOS;
  uint32_t p = e & (B - 1u);
  const bool act = i < count && p < count;  // (p < count always with a correct sort; keeps every read in the block)
  if (!act) p = 0;
  const uint32_t xe = xact ? sb[i0 + lane - MX] : NOPOS;
  const uint32_t xp = xe & (B - 1u);
  lent[MX + lane] = act ? e : NOPOS;
  lneed[MX + lane] = 0;
  if (lane < MX) {
    lent[lane] = xact && xp < count ? xe : NOPOS;
    lneed[lane] = 0;
  }
  wave_fence();
  // which of the slots before this one are candidates (bit k - 1 for slot i - k), from the sorted entries alone: a
  // slot's tag is its key's bits below the bucket's, so the low 12 bits of the tags are those of the keys
  const uint32_t tag = e >> blog;
  uint32_t okm = 0;
#pragma unroll
  for (uint32_t k = 1; k <= MX; ++k) {
    if (k > ncand) break;  // (uniform)
    const uint32_t ce = lent[MX + lane - k];
    const bool ok = act && ce != NOPOS && (((ce >> blog) ^ tag) & 0xFFFu) == 0 && (ce & (B - 1u)) < p;
    okm |= ok ? 1u << (k - 1) : 0u;
    if (ok) lneed[MX + lane - k] = 1u;  // (lanes that share a candidate write the same value)
  }
  wave_fence();
  // windows only where they are compared: a lane with a candidate, a slot that is one (about 7 in 10 of the
  // positions of LZ-like data; the others' gathers are not made)
  const bool want = act && (okm != 0u || lneed[MX + lane] != 0u);
  const bool xwant = lane < MX && lneed[lane] != 0u;
  uint32_t w[NW], xw[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) w[q] = xw[q] = 0;
  if (want) load_window<NW>(inb, n, p, w);
  if (xwant) load_window<NW>(inb, n, xp < count ? xp : 0u, xw);
  // (the byte before p is read only where the window starts with 4 equal bytes: those lanes alone take the load)
  bool run = false;
  if (okm && p > 0 && w[0] == (w[0] & 0xFFu) * 0x01010101u) run = inb[p - 1] == (w[0] & 0xFFu);
  if (want) {
#pragma unroll
    for (int q = 0; q < NW; ++q) lwin[(MX + lane) * NW + q] = w[q];
  }
  if (xwant) {
#pragma unroll
    for (int q = 0; q < NW; ++q) lwin[lane * NW + q] = xw[q];
  }
  wave_fence();
  const uint32_t maxlen = min(lencap, n - p);
  // best starts at 3, so that "longer than best" also means "at least 4 long" (a match is)
  uint32_t best = 3, bestd = 0;
  if (run && okm) {  // a run of one byte value (kernels/zero.cl:219-222): 5 equal bytes from p - 1
    uint32_t r[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) r[q] = w[0];
    const uint32_t l = min(common_prefix<NW>(w, r), maxlen);
    if (l > best) { best = l; bestd = 1; }
  }
  // every candidate compared, with selects rather than branches (lanes differ in how many candidates they have)
#pragma unroll
  for (uint32_t k = 1; k <= MX; ++k) {
    if (k > ncand) break;  // (uniform)
    const uint32_t idx = MX + lane - k;
    const uint32_t cp = lent[idx] & (B - 1u);
    uint32_t c[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) c[q] = lwin[idx * NW + q];
    const uint32_t l = min(common_prefix<NW>(w, c), maxlen);
    const bool better = ((okm >> (k - 1)) & 1u) && l > best;
    best = better ? l : best;
    bestd = better ? p - cp : bestd;
  }
  if (act && best >= 4) mb[p] = (best << 20) | (bestd - 1);  // (the others stay 0: the array was cleared)
}

// 256 sorted slots of block b, from slot s0 (a wavefront takes 64 of them)
__device__ __forceinline__ void enc_match_slots(const Block* __restrict__ blocks, const uint8_t* __restrict__ in, uint32_t blog,
                                       uint32_t ncand, uint32_t lencap,
                                       const uint32_t* __restrict__ srt, uint32_t* __restrict__ match, uint32_t b,
                                       uint32_t s0, uint32_t* sh) {  // sh: MATCH_LDS_WORDS per wavefront
  const uint32_t i0 = s0 + (threadIdx.x & ~63u);
  const Block blk = blocks[b];
  const uint32_t n = enc_len(blk, blog);
  if (i0 >= n) return;  // (uniform over the wavefront)
  if (blk.level) level_params(blk.level, ncand, lencap);
  const uint32_t count = n >= 4 ? n - 3 : 0;
  const uint8_t* inb = in + blk.src;
  const uint32_t* sb = srt + ((size_t)b << blog);
  uint32_t* mb = match + ((size_t)b << blog);
  if (lencap <= 16)
    match_slots<4>(inb, n, count, blog, min(ncand, MX), lencap, sb, i0, mb, sh + (threadIdx.x >> 6) * MATCH_LDS_WORDS);
  else
    match_slots<8>(inb, n, count, blog, min(ncand, MX), min(lencap, 32u), sb, i0, mb,
                   sh + (threadIdx.x >> 6) * MATCH_LDS_WORDS);
}
__global__ __launch_bounds__(256) void k_enc_match(const Block* __restrict__ blocks, const uint8_t* __restrict__ in,
                                                   uint32_t blog, uint32_t ncand, uint32_t lencap,
                                                   const uint32_t* __restrict__ srt, uint32_t* __restrict__ match) {
  __shared__ uint32_t sh[4 * MATCH_LDS_WORDS];
  const uint32_t wg = xcd_swizzle(blockIdx.x, gridDim.x);
  enc_match_slots(blocks, in, blog, ncand, lencap, srt, match, wg >> (blog - 8),
                  (wg << 8) & ((1u << blog) - 1u), sh);
}

// ---------------------------------------------------------------- 4. lazy matching (kernels/zero.cl:239-255)
// A match gives way to a better one a byte later: the value the parse sees at p, from the match finder's values at
// p and p + 1 (0 past the block). It is computed where it is used, by the parse and the token pass. The parse
// writes a match it finishes past LENCAP back into the match array; lengths here are taken at most LENCAP, and a
// finished match was LENCAP long before, so the decisions come out the same before and after.
__device__ inline uint32_t lazy_of(uint32_t v, uint32_t v2, uint32_t lencap) {
  const uint32_t l = min(v >> 20, lencap), l2 = min(v2 >> 20, lencap);
  if (l && l2) {
    const int c = (int)bitlen((v & 0xFFFFFu) + 1u), c2 = (int)bitlen((v2 & 0xFFFFFu) + 1u);
    if (8 * (int)l2 - c2 - 8 > 8 * (int)l - c) return 0u;
  }
  return v;
}

// ---------------------------------------------------------------- 5. the greedy parse (one workgroup per block)
// the length of the match a value holds at p, or 0 when it holds none (a value that could not be the match finder's
// is taken as none: a bug elsewhere then costs ratio, never an access outside the block)
__device__ inline uint32_t match_at(uint32_t v, uint32_t p, uint32_t n) {
  const uint32_t l = v >> 20, d = (v & 0xFFFFFu) + 1u;
  return l >= 4 && l <= n - p && d <= p ? l : 0u;
}

// prefix_len for long matches: 64 bytes a step, all 8 loads of a step in flight together
__device__ inline uint32_t prefix_len_long(const uint8_t* inb, uint32_t a, uint32_t b, uint32_t maxlen) {
  uint32_t l = 0;
  while (l + 64 <= maxlen) {
    uint64_t x[8], y[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; u += 2) {
      __builtin_memcpy(&x[u], inb + a + l + 8 * u, 16);
      __builtin_memcpy(&y[u], inb + b + l + 8 * u, 16);
    }
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u)
      if (x[u] != y[u]) return l + 8 * u + ((uint32_t)__builtin_ctzll(x[u] ^ y[u]) >> 3);
    l += 64;
  }
  return l + prefix_len(inb, a + l, b + l, maxlen - l);
}

// the position after the token at p; a match cut at LENCAP is finished here, up to 4095 bytes (kernels/zero.cl:259-271)
__device__ inline uint32_t nextpos_v(uint32_t* m, const uint8_t* inb, uint32_t n, uint32_t lencap, uint32_t p,
                                     uint32_t v) {  // v = m[p]
  uint32_t l = match_at(v, p, n);
  if (l == lencap) {
    const uint32_t maxlen = min(4095u, n - p);
    if (maxlen > l) {
      const uint32_t j = p - (v & 0xFFFFFu) - 1u;
      const uint32_t x = l + prefix_len_long(inb, j + l, p + l, maxlen - l);
      if (x != l) { m[p] = (x << 20) | (v & 0xFFFFFu); l = x; }
    }
  }
  return p + (l ? l : 1u);
}
__device__ inline uint32_t nextpos(uint32_t* m, const uint8_t* inb, uint32_t n, uint32_t lencap, uint32_t p) {
  return nextpos_v(m, inb, n, lencap, p, lazy_of(m[p], p + 1 < n ? m[p + 1] : 0u, lencap));
}

// One thread's marks in a block's chain bitmap, gathered a word at a time in a register and stored as the words are
// passed, g (1, 2 or 4) aligned words to a store: the thread owns the words it writes (a segment's, a multiple of g
// words from a multiple of g) and stores each once, in order, zero words included. Fewer stores matter beyond
// their own cost: a load waits for every memory access issued before it, stores included.
struct MarkWriter {
  uint32_t* mk;  // the block's bitmap
  uint32_t cw;   // the word being gathered
  uint32_t bits;
  uint32_t g;
  uint32_t w0 = 0, w1 = 0, w2 = 0;  // the group's words before cw (the last in w2)
  __device__ void next() {
    if (g == 4u && (cw & 3u) == 3u) {
      *reinterpret_cast<uint4*>(mk + cw - 3u) = make_uint4(w0, w1, w2, bits);
    } else if (g == 2u && (cw & 1u) == 1u) {
      *reinterpret_cast<uint2*>(mk + cw - 1u) = make_uint2(w2, bits);
    } else if (g == 1u) {
      mk[cw] = bits;
    }
    w0 = w1; w1 = w2; w2 = bits;
    bits = 0;
    ++cw;
  }
  // stores the words before position a's: the gathered word and the rest of its group, then whole groups of zero
  // words (a long match skips many), then the zero words before a's in its group
  __device__ void to(uint32_t a) {
    const uint32_t t = a >> 5;
    if (t <= cw) return;
    do next(); while (cw < t && (cw & (g - 1u)));
    for (; cw + g <= t; cw += g) {
      if (g == 4u) *reinterpret_cast<uint4*>(mk + cw) = make_uint4(0u, 0u, 0u, 0u);
      else if (g == 2u) *reinterpret_cast<uint2*>(mk + cw) = make_uint2(0u, 0u);
      else mk[cw] = 0u;
    }
    while (cw < t) next();
  }
  // sets the bits a .. b - 1 (a < b <= a + 32, a at or past the gathered word's start)
  __device__ void range(uint32_t a, uint32_t b) {
    to(a);
    const uint64_t r = ((1ull << (b - a)) - 1ull) << (a & 31u);
    bits |= (uint32_t)r;
    if ((uint32_t)(r >> 32)) { next(); bits = (uint32_t)(r >> 32); }
  }
};

// the chain from e through the positions before send: where it meets a mark in mk, that position | 1 << 31; else
// the first position at or past send. The mark word and the matches at t are requested together.
__device__ inline uint32_t join_walk(uint32_t* m, const uint8_t* inb, uint32_t n, uint32_t lencap, const uint32_t* mk,
                                     uint32_t e, uint32_t send) {
  uint32_t t = e;
  while (t < send) {
    const uint32_t w = mk[t >> 5], v = m[t], v2 = t + 1 < n ? m[t + 1] : 0u;
    if ((w >> (t & 31u)) & 1u) return t | 0x80000000u;
    t = nextpos_v(m, inb, n, lencap, t, lazy_of(v, v2, lencap));
  }
  return t;
}

// Each of the first `walkers` threads (a power of two, 32 <= B / walkers <= 1024) walks its segment of B / walkers positions
// from the segment's start, marking token starts in the block's chain bitmap mark[b] (B bits; each walker writes its
// own segment's words); the join (below) puts the true chain in their place: where the chain meets a walk the two
// coincide from there on (the chain is a pure function of the position), so only the walk's prefix is wrong; where it
// never meets one, the walk is wrong whole. The workgroup's barriers order one thread's bitmap stores before another's
// loads.
__device__ __forceinline__ void enc_parse_block(const Block* __restrict__ blocks, const uint8_t* __restrict__ in, uint32_t blog,
                                       uint32_t walkers, uint32_t lencap, uint32_t* __restrict__ match,
                                       uint32_t* __restrict__ mark, uint32_t b) {
  extern __shared__ uint32_t sm[];  // ex, sin, sout (walkers words each), windows
  const uint32_t B = 1u << blog, SEG = B / walkers;  // (blockDim.x >= walkers)
  const uint32_t k = threadIdx.x;
  uint32_t* ex = sm;
  uint32_t* sin = ex + walkers;    // a segment's entry as walked
  uint32_t* sout = sin + walkers;  // where that walk met walk q (| 1 << 31), or left the segment
  const Block blk = blocks[b];
  const uint8_t* inb = in + blk.src;
  const uint32_t n = enc_len(blk, blog);
  uint32_t* m = match + ((size_t)b << blog);
  uint32_t* mk = mark + ((size_t)b << (blog - 5));
  if (blk.level) { uint32_t nc; level_params(blk.level, nc, lencap); }
  if (k < walkers) {
    // m is read PARSE_WIN entries at a time into this walker's LDS window: the many walks in flight would otherwise
    // fetch a line of m for each step, their lines leaving the caches before the next step comes. A step goes from p straight to the window's next position with a match
    // (the positions before it are literals, one chain position each) or, with none, to the window's end.
    uint32_t* win = sout + walkers + (PARSE_WIN + 1u) * k;
    const uint32_t s = k * SEG, send = min(s + SEG, n);
    MarkWriter mw{mk, s >> 5, 0u, SEG >= 128u ? 4u : SEG >= 64u ? 2u : 1u};
    uint32_t p = s, wb = s - PARSE_WIN, mm = 0;  // window base (none yet), which of its positions have a match
    while (p < send) {
      if (p - wb >= PARSE_WIN) {
        wb = p & ~(PARSE_WIN - 1u);
        uint32_t v[PARSE_WIN + 1];
        __builtin_memcpy(v, m + wb, 4 * PARSE_WIN);
        v[PARSE_WIN] = m[min(wb + PARSE_WIN, B - 1u)];  // (used only when wb + PARSE_WIN < n)
        mm = 0;
#pragma unroll
        for (uint32_t q = 0; q < PARSE_WIN; ++q) {
          const uint32_t lz = lazy_of(v[q], wb + q + 1u < n ? v[q + 1] : 0u, lencap);
          win[q] = lz;
          mm |= (match_at(lz, wb + q, n) ? 1u : 0u) << q;
        }
      }
      const uint32_t ahead = mm & (~0u << (p - wb));
      const uint32_t q = ahead ? wb + (uint32_t)__builtin_ctz(ahead) : wb + PARSE_WIN;
      if (q >= send) {  // literals to the segment's end
        mw.range(p, send);
        p = send;
      } else if (!ahead) {
        mw.range(p, q);
        p = q;
      } else {
        mw.range(p, q + 1);  // (a segment owns whole words: SEG is a multiple of 32)
        p = nextpos_v(m, inb, n, lencap, q, win[q - wb]);
      }
    }
    mw.to(s + SEG);  // the rest of the segment's words (zero past the block's end)
    ex[k] = s < n ? p : s;
  }
  __syncthreads();
  // the join. Segment q's true entry is where the chain of the segments before it leaves them; from there the chain
  // runs until it meets a mark of walk q (from then on the two coincide, and the chain leaves the segment at ex[q])
  // or leaves the segment. Each segment first assumes its entry to be ex[q - 1] (right whenever the chain met walk
  // q - 1) and walks from it, all segments at once; thread 0 then follows the entries in order and walks again only
  // where an assumption was wrong; last, each segment clears the walk's marks before the meeting point and marks the
  // chain's own positions there.
  const uint32_t nseg = min(walkers, (n + SEG - 1) / SEG);
  if (k < nseg) {
    const uint32_t e = k ? ex[k - 1] : 0u;
    sin[k] = e;
    sout[k] = join_walk(m, inb, n, lencap, mk, e, min(k * SEG + SEG, n));
  }
  __syncthreads();
  if (k == 0) {
    uint32_t e = 0;
    for (uint32_t q = 0; q < nseg; ++q) {
      uint32_t r = sout[q];
      if (sin[q] != e) {
        r = join_walk(m, inb, n, lencap, mk, e, min(q * SEG + SEG, n));
        sin[q] = e;
        sout[q] = r;
      }
      e = r >> 31 ? ex[q] : r;
    }
  }
  __syncthreads();
  if (k < nseg) {
    // the chain's positions replace the walk's marks from the segment's start up to the meeting point: whole words
    // stored anew, the word holding the meeting point merged with the walk's marks from there on
    const uint32_t s = k * SEG, send = min(s + SEG, n), r = sout[k];
    const uint32_t clear_to = r >> 31 ? r & 0x7FFFFFFFu : send;
    MarkWriter mw{mk, s >> 5, 0u, 1u};
    for (uint32_t p = sin[k]; p < clear_to; p = nextpos(m, inb, n, lencap, p)) {
      mw.to(p);
      mw.bits |= 1u << (p & 31u);
    }
    mw.to(clear_to);
    if (clear_to & 31u) mk[mw.cw] = mw.bits | (mk[mw.cw] & (~0u << (clear_to & 31u)));
  }
}
// at most 64 VGPRs (8 wavefronts per SIMD): two workgroups of 1024 walkers (2^20-byte blocks) then fit a CU, not one
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_enc_parse(
    const Block* __restrict__ blocks, const uint8_t* __restrict__ in, uint32_t blog, uint32_t walkers,
    uint32_t lencap, uint32_t* __restrict__ match, uint32_t* __restrict__ mark) {
  enc_parse_block(blocks, in, blog, walkers, lencap, match, mark, blockIdx.x);
}

// ---------------------------------------------------------------- 6. tokens and streams (one wavefront per block)
// From the chain: lit (the literal bytes), then per token (in order) its literal run L, length M, distance D, and
// the groups (src/codec2.h:102-136; kernels/zero.cl:375-451): a group closes at 255 tokens, or before a match whose
// source reaches into the group (reading only its own literals is allowed). sym is the offset's bit length, or
// 250..252 for one of the last three offsets as of the end of the group before.
// cnt[b*8 + 0..4] = tokens, literals, ext bytes, offset-bit bytes, groups.
struct TokBufs {
  uint8_t* lit;      // B per block
  uint32_t* LB;      // literals before each token (TCAP per block)
  uint32_t* TM;      // match length (TCAP)
  uint32_t* TD;      // distance (TCAP)
  uint8_t* tok;      // TCAP
  uint8_t* sym;      // TCAP
  uint8_t* grp;      // TCAP
  uint8_t* ext;      // ECAP
  uint32_t* offw;    // OCAP/4 + 1 words
  uint32_t* cnt;     // 8 per block
};

__device__ __forceinline__ void enc_tokens_block(const Block* __restrict__ blocks, const uint8_t* __restrict__ in,
                                        uint32_t blog, uint32_t lencap, const uint32_t* __restrict__ match,
                                        const uint32_t* __restrict__ mark, TokBufs T, uint32_t b) {
  __shared__ uint32_t stage[48];
  const Geo G = geo_of(blog);
  const uint32_t lane = threadIdx.x;
  const Block blk = blocks[b];
  const uint8_t* inb = in + blk.src;
  const uint32_t n = enc_len(blk, blog);
  const uint32_t* m = match + ((size_t)b << blog);
  if (blk.level) { uint32_t nc; level_params(blk.level, nc, lencap); }
  const uint32_t* mk = mark + ((size_t)b << (blog - 5));
  uint8_t* lit = T.lit + ((size_t)b << blog);
  uint32_t* LB = T.LB + (size_t)b * G.TCAP;
  uint32_t* TM = T.TM + (size_t)b * G.TCAP;
  uint32_t* TD = T.TD + (size_t)b * G.TCAP;
  // phase A: the chain, 512 positions a step, 8 consecutive ones per lane (ranks from one scan per step); every load
  // of the next step is in flight while this one is written
  uint32_t nt = 0, nl = 0;
  const bool big = n >= 8;
  uint32_t cm[8], nm[8];
  uint64_t cx, nx = 0;
  uint32_t cw, nw = 0;
  __builtin_memcpy(cm, m + 8 * lane, 32);
  cx = big ? ld8z(inb, n, 8 * lane) : ld8z_small(inb, n, 8 * lane);
  cw = mk[lane >> 2];
  for (uint32_t p0 = 0; p0 < n; p0 += 512) {
    const uint32_t qn = p0 + 512 + 8 * lane;
    if (p0 + 512 < n) {  // (uniform)
      __builtin_memcpy(nm, m + qn, 32);
      nx = ld8z(inb, n, qn);
      nw = mk[qn >> 5];
    }
    const uint32_t base = p0 + 8 * lane;
    const uint32_t bits = (cw >> (base & 31u)) & 0xFFu;  // which of this lane's 8 positions are on the chain
    // the lazy values: the match value after the lane's last comes from the next lane, or the next step's lane 0
    // (the lane move runs with every lane active: a lane that sits out gives the lane before it nothing)
    const uint32_t from_next = from_next_lane(cm[0]);
    const uint32_t next_step = p0 + 512 < n ? readlane(nm[0], 0) : 0u;
    const uint32_t after = lane == 63u ? next_step : from_next;
    uint32_t tmask = 0;
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      cm[u] = lazy_of(cm[u], base + u + 1u < n ? (u < 7 ? cm[u + 1] : after) : 0u, lencap);
      if (((bits >> u) & 1u) && match_at(cm[u], base + u, n)) tmask |= 1u << u;
    }
    const uint32_t lmask = bits & ~tmask;
    const uint32_t lc = (uint32_t)__builtin_popcount(lmask), tc = (uint32_t)__builtin_popcount(tmask);
    const uint32_t li = wave_scan_incl(lc), ti = wave_scan_incl(tc);
    uint32_t lq = nl + li - lc, tq = nt + ti - tc;
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      if ((lmask >> u) & 1u) lit[lq++] = (uint8_t)(cx >> (8u * u));
      if ((tmask >> u) & 1u) {
        LB[tq] = lq; TM[tq] = cm[u] >> 20; TD[tq] = (cm[u] & 0xFFFFFu) + 1u;
        ++tq;
      }
    }
    nl += readlane(li, 63);
    nt += readlane(ti, 63);
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) cm[u] = nm[u];
    cx = nx;
    cw = nw;
  }
  wave_fence();
  // phase B: 64 tokens at a time
  uint8_t* tokb = T.tok + (size_t)b * G.TCAP;
  uint8_t* symb = T.sym + (size_t)b * G.TCAP;
  uint8_t* grpb = T.grp + (size_t)b * G.TCAP;
  uint8_t* extb = T.ext + (size_t)b * G.ECAP;
  uint32_t* ow = T.offw + (size_t)b * (G.OCAP / 4u + 1u);
  uint32_t op = 0, ei = 0, bp = 0, ng = 0, gstart = 0, gt = 0, carry = 0;
  uint32_t r0 = 1, r1 = 4, r2 = 8, l0 = 1, l1 = 4, l2 = 8;
  // (each chunk's three loads are issued one chunk ahead, at clamped indexes)
  uint32_t lbr = 0, Mr = 0, Dr = 1, lbprev = 0;
  if (nt) { const uint32_t q = min(lane, nt - 1); lbr = LB[q]; Mr = TM[q]; Dr = TD[q]; }
  for (uint32_t t = 0; t < nt; t += 64) {
    const uint32_t c = min(64u, nt - t);
    const bool on = lane < c;
    const uint32_t q = t + lane;
    const uint32_t qn = min(t + 64 + lane, nt - 1);
    const uint32_t lbn = LB[qn], Mn = TM[qn], Dn = TD[qn];
    const uint32_t lb = on ? lbr : 0u;
    const uint32_t up = from_prev_lane(lbr);
    const uint32_t lprev = on ? (lane ? up : lbprev) : 0u;
    const uint32_t L = lb - lprev;
    const uint32_t M = on ? Mr : 0u;
    const uint32_t D = on ? Dr : 1u;
    lbprev = readlane(lbr, 63);
    lbr = lbn; Mr = Mn; Dr = Dn;
    const uint32_t opk = op + wave_scan_excl(L + M);
    const uint32_t i = opk + L, j = i - D;
    const uint32_t need = j < opk ? j + min(M, D) : 0u;  // how far the source reaches, unless its own literals
    uint32_t sv = 0;
    uint32_t lo = 0;
    while (lo < c) {
      const bool viol = on && lane >= lo && ((need > gstart && !(lane == lo && gt == 0)) || gt + (lane - lo) >= 255u);
      const uint64_t vm = ballot(viol);
      const uint32_t first = vm ? min((uint32_t)__builtin_ctzll(vm), c) : c;
      if (lane >= lo && lane < first) sv = D == r0 ? 250u : D == r1 ? 251u : D == r2 ? 252u : bitlen(D);
      for (uint32_t x = first > lo + 3 ? first - 3 : lo; x < first; ++x) {
        const uint32_t v = readlane(D, x);
        l2 = l1; l1 = l0; l0 = v;
      }
      gt += first - lo;
      if (first < c) {
        if (gt) { if (lane == 0) grpb[ng] = (uint8_t)gt; ++ng; }
        gt = 0;
        gstart = readlane(opk, first);
        r0 = l0; r1 = l1; r2 = l2;
      }
      lo = first;
    }
    if (on) {
      symb[q] = (uint8_t)sv;
      tokb[q] = (uint8_t)((min(L, 15u) << 4) | min(M - 4u, 15u));
    }
    const uint32_t fl = on && L >= 15u, fm = on && M >= 19u;
    const uint32_t xq = ei + wave_scan_excl(3u * fl + 2u * fm);
    if (fl) { const uint32_t x = L - 15u; extb[xq] = (uint8_t)x; extb[xq + 1] = (uint8_t)(x >> 8); extb[xq + 2] = (uint8_t)(x >> 16); }
    if (fm) { const uint32_t x = M - 19u, y = xq + 3u * fl; extb[y] = (uint8_t)x; extb[y + 1] = (uint8_t)(x >> 8); }
    // offset bits, LSB first: gathered in LDS words, then stored whole (the last one carried into the next chunk)
    const uint32_t nbb = on && sv < 250u ? sv : 0u;
    const uint32_t bq = bp + wave_scan_excl(nbb);
    const uint32_t tb = wave_sum(nbb);
    const uint32_t wb = bp >> 5;
    if (lane < 48) stage[lane] = 0;
    __syncthreads();
    if (nbb) {
      const uint32_t extra = D - (1u << nbb);
      const uint32_t w = (bq >> 5) - wb, sh = bq & 31u;
      atomicOr(&stage[w], extra << sh);
      if (sh + nbb > 32u) atomicOr(&stage[w + 1], extra >> (32u - sh));
    }
    __syncthreads();
    const uint32_t bend = bp + tb;
    const uint32_t nw = (bend >> 5) - wb + 1;
    if (lane < nw) ow[wb + lane] = stage[lane] | (lane == 0 ? carry : 0u);
    carry = stage[(bend >> 5) - wb] | ((bend >> 5) == wb ? carry : 0u);
    __syncthreads();
    op += wave_sum(L + M);
    ei += wave_sum(3u * fl + 2u * fm);
    bp = bend;
  }
  if (gt) { if (lane == 0) grpb[ng] = (uint8_t)gt; ++ng; }
  if (lane == 0) {
    uint32_t* cn = T.cnt + (size_t)b * 8;
    cn[0] = nt; cn[1] = nl; cn[2] = ei; cn[3] = (bp + 7u) >> 3; cn[4] = ng;
  }
}
__global__ __launch_bounds__(64) void k_enc_tokens(const Block* __restrict__ blocks, const uint8_t* __restrict__ in,
                                                   uint32_t blog, uint32_t lencap,
                                                   const uint32_t* __restrict__ match,
                                                   const uint32_t* __restrict__ mark, TokBufs T) {
  enc_tokens_block(blocks, in, blog, lencap, match, mark, T, blockIdx.x);
}

// ---------------------------------------------------------------- 7. rANS per (block, stream)
constexpr uint32_t RING_STRIDE = 144;  // a lane's LDS ring of 128 bytes, 8 of spill, padded to 16
// 64 bytes from LDS (16-byte aligned) to memory (64-byte aligned)
__device__ inline void copy64(uint8_t* dst, const uint8_t* src) {
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
  for (int q = 0; q < 4; ++q) d[q] = s[q];
}
struct RansBufs {
  uint8_t* lanes;     // lane scratch, LANES bytes per block
  uint16_t* freq;     // 256 per (block, stream)
  uint32_t* lanelen;  // 64 per (block, stream)
  uint32_t* sinfo;    // 4 per (block, stream): kind, section length, present symbols, symbols
};
__host__ __device__ inline uint32_t lane_cap(const Geo& G, uint32_t s) {
  const uint32_t maxn = s == 0 ? G.B : s == 2 ? G.ECAP : G.TCAP;
  // at most 2 bytes per symbol, plus the 4-byte state; a multiple of 64 (lanes leave in aligned 64-byte pieces)
  return (2u * ((maxn + 63u) / 64u) + 16u + 63u) & ~63u;
}
__host__ __device__ inline uint32_t lane_off(const Geo& G, uint32_t s) {
  uint32_t o = 0;
  for (uint32_t i = 0; i < s; ++i) o += 64u * lane_cap(G, i);
  return o;
}
__host__ __device__ inline uint32_t lanes_per_block(const Geo& G) { return lane_off(G, 5); }

__device__ inline const uint8_t* stream_syms(const TokBufs& T, const Geo& G, uint32_t b, uint32_t s) {
  switch (s) {
    case 0: return T.lit + ((size_t)b << G.blog);
    case 1: return T.tok + (size_t)b * G.TCAP;
    case 2: return T.ext + (size_t)b * G.ECAP;
    case 3: return T.sym + (size_t)b * G.TCAP;
    default: return T.grp + (size_t)b * G.TCAP;
  }
}
__device__ inline uint32_t stream_count(const uint32_t* cn, uint32_t s) {
  return s == 0 ? cn[1] : s == 2 ? cn[2] : s == 4 ? cn[4] : cn[0];
}

// stream s of block b (g = 5 b + s indexes the outputs)
__device__ __forceinline__ void enc_rans_stream(uint32_t blog, TokBufs T, RansBufs R, uint32_t b, uint32_t s) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t fc[256];   // freq | cum << 16
  __shared__ uint32_t rcp[256];  // ceil(2^(31+s) / freq), s = ceil(log2 freq)
  const Geo G = geo_of(blog);
  const uint32_t g = 5u * b + s, lane = threadIdx.x;
  const uint32_t* cn = T.cnt + (size_t)b * 8;
  const uint32_t n = stream_count(cn, s);
  const uint8_t* sy = stream_syms(T, G, b, s);
  uint32_t* si = R.sinfo + (size_t)g * 4;
  if (n == 0) {
    if (lane == 0) { si[0] = 0; si[1] = 0; si[2] = 0; si[3] = 0; }
    return;
  }
  for (uint32_t i = lane; i < 256; i += 64) hist[i] = 0;
  __syncthreads();
  {  // 16 bytes per lane per round of 1 KiB, the loads of 4 rounds requested together; the last partial round
     // byte by byte (loads at any alignment: the ext stream's start is not 16-byte aligned)
    const uint32_t nfull = n & ~1023u;
    for (uint32_t base = 0; base < nfull; base += 4096u) {
      uint32_t v[4][4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        if (base + 1024u * k < nfull) __builtin_memcpy(v[k], sy + base + 1024u * k + 16u * lane, 16);
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        if (base + 1024u * k < nfull) {
#pragma unroll
          for (uint32_t w = 0; w < 4; ++w)
#pragma unroll
            for (uint32_t by = 0; by < 4; ++by) atomicAdd(&hist[(v[k][w] >> (8u * by)) & 255u], 1u);
        }
      }
    }
    for (uint32_t i = nfull + lane; i < n; i += 64) atomicAdd(&hist[sy[i]], 1u);
  }
  __syncthreads();
  // 12-bit frequencies exactly as zb::normalize (src/codec.h:41-66), lane owns symbols 4*lane..4*lane+3
  uint32_t f[4], sum = 0, npres = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t c = hist[lane * 4 + k];
    f[k] = 0;
    if (c) { f[k] = (uint32_t)(((uint64_t)c << 12) / n); if (!f[k]) f[k] = 1; ++npres; }
    sum += f[k];
  }
  sum = wave_sum(sum);
  npres = wave_sum(npres);
  if (sum > 4096u) {
    // The reference takes one from the first largest frequency (> 1) per step. After E = sum - 4096 steps every
    // frequency above T+1 is down to T+1, and the first E' of those at T+1 (by symbol) are down to T, where T is
    // the largest level with need(T) = sum of max(0, f - T) >= E and E' = E - need(T+1).
    const uint32_t E = sum - 4096u;
    uint32_t lo = 1, hi = 4096;  // need(lo) >= E holds at lo = 1 (since present <= 4096)
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      uint32_t nd = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) nd += f[k] > mid ? f[k] - mid : 0u;
      nd = wave_sum(nd);
      if (nd >= E) lo = mid; else hi = mid - 1;
    }
    const uint32_t T1 = lo + 1;
    uint32_t nd1 = 0, at = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) { nd1 += f[k] > T1 ? f[k] - T1 : 0u; at += f[k] >= T1 ? 1u : 0u; }
    nd1 = wave_sum(nd1);
    const uint32_t E2 = E - nd1;
    uint32_t rank = wave_scan_excl(at);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (f[k] >= T1) { f[k] = rank < E2 ? T1 - 1 : T1; ++rank; }
  } else if (sum < 4096u) {
    // the deficit to the first largest
    uint32_t key = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) key = max(key, f[k] << 8 | (255u - (lane * 4 + k)));
    key = wave_max(key);
    const uint32_t best = 255u - (key & 255u);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (lane * 4 + k == best) f[k] += 4096u - sum;
  }
  uint32_t fs = f[0] + f[1] + f[2] + f[3];
  uint32_t c = wave_scan_excl(fs);
  uint16_t* fo = R.freq + (size_t)g * 256;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t sym = lane * 4 + k;
    fc[sym] = f[k] | c << 16;
    fo[sym] = (uint16_t)f[k];
    if (f[k]) {
      const uint32_t sh = f[k] > 1 ? 32u - (uint32_t)__builtin_clz(f[k] - 1) : 0u;
      rcp[sym] = (uint32_t)((((uint64_t)1 << (31 + sh)) + f[k] - 1) / f[k]);
    }
    c += f[k];
  }
  __syncthreads();
  // lane j encodes symbols j, j+64, ... from the last one back into its scratch, backwards (src/codec.h:112-126).
  // Steps go 4 at a time (a group). The 1024 symbols of 4 groups arrive in LDS from one coalesced 16-byte load per
  // lane made 4 groups earlier (the symbol streams keep 1 KiB of slack for it). The (at most 8) bytes the lane emits
  // in a group go, as one 8-byte write ending where its bytes stand (the bytes below them are garbage that the next
  // group's write covers), into the lane's ring of 128 bytes in LDS. Every 4 groups (at most 32 bytes) a lane whose
  // next aligned 64-byte piece is complete stores it whole: no partial line of the scratch goes to memory. Those
  // stores are made at the start of the next 4 groups, after its symbols are in LDS and before its load, so that a
  // wait for a load never waits for stores issued after it. (Between two such points the lane holds at most 63 + 32
  // bytes not yet stored, and writes 8 below them: the ring never overtakes them.)
  __shared__ uint32_t stage[256];
  __shared__ __align__(16) uint8_t rings[64 * RING_STRIDE];
  uint8_t* rl = rings + lane * RING_STRIDE;  // byte at address a: rl[(a - end) mod 128], with 8 bytes of spill past 128
  const uint32_t LC = lane_cap(G, s);
  uint8_t* end = R.lanes + (size_t)b * lanes_per_block(G) + lane_off(G, s) + (size_t)(lane + 1) * LC;
  uint8_t* ptr = end;
  uint8_t* fl = end;     // [fl, end) is in memory; fl is 64-byte aligned
  uint32_t ro = 120;     // ring offset of ptr - 8
  uint32_t fro = 64;     // ring offset of fl - 64
  uint32_t x = 1u << 23;
  const uint32_t NT = (n + 63u) / 64u;                 // steps: symbol lane + 64 t at step t, t from NT-1 down
  int32_t J = (int32_t)((NT - 1u) >> 4);               // groups [4J, 4J + 3], steps [16J, 16J + 15]
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  v4u nxt;
  __builtin_memcpy(&nxt, sy + 1024u * (uint32_t)J + 16u * lane, 16);
  bool pend = false;  // a piece of the last 4 groups is to be stored
  for (; J >= 0; --J) {
    wave_fence();
    __builtin_memcpy(stage + 4u * lane, &nxt, 16);
    wave_fence();
    if (pend) {
      copy64(fl - 64, rl + fro);
      fl -= 64;
      fro = (fro - 64u) & 127u;
    }
    __builtin_memcpy(&nxt, sy + 1024u * (uint32_t)(J > 0 ? J - 1 : 0) + 16u * lane, 16);  // (no branch: at J = 0
                                                                                         // an unused reload)
    const uint8_t* st8 = (const uint8_t*)stage;
#pragma unroll 1
    for (int32_t jj = 3; jj >= 0; --jj) {
      const int32_t j = 4 * J + jj;  // (groups past the last step change nothing: no step of theirs is valid)
      uint64_t acc = 0;
      uint32_t k = 0;  // bytes emitted in this group, the first at the top of acc
#pragma unroll
      for (int q = 3; q >= 0; --q) {
        const uint32_t t = 4u * (uint32_t)j + (uint32_t)q;
        const bool valid = t < NT && lane + 64u * t < n;
        const uint32_t sym = st8[256u * (uint32_t)jj + 64u * (uint32_t)q + lane];
        const uint32_t e = fc[sym];
        const uint32_t fq = e & 0xFFFFu;
        const uint32_t xmax = fq << 19;
        // at most two bytes leave (x < 2^31, and x < 2^23 < ... after one; fq = 1 may take a second)
        const bool e1 = valid && x >= xmax;
        const uint32_t x1 = e1 ? x >> 8 : x;
        const bool e2 = e1 && x1 >= xmax;
        const uint32_t x2 = e2 ? x1 >> 8 : x1;
        acc |= e1 ? (uint64_t)(x & 255u) << (8u * (7u - k)) : 0ull;
        k += e1 ? 1u : 0u;
        acc |= e2 ? (uint64_t)(x1 & 255u) << (8u * (7u - k)) : 0ull;
        k += e2 ? 1u : 0u;
        // exact x / fq for x < 2^31: floor(x * ceil(2^(31+s)/fq) / 2^(31+s)) with 2^s >= fq
        const uint32_t sh = fq > 1 ? 32u - (uint32_t)__builtin_clz(fq - 1) : 0u;
        const uint32_t qt = (uint32_t)(((uint64_t)x2 * rcp[sym]) >> (31 + sh));
        const uint32_t xn = (qt << 12) + (x2 - qt * fq) + (e >> 16);
        x = valid ? xn : x;
      }
      __builtin_memcpy(rl + ro, &acc, 8);
      if (ro > 120u) {  // (the part past the ring's end belongs at its start)
        const uint32_t over = ro - 120u;  // bytes past 128: 1..7
        const uint64_t hi = acc >> (8u * (8u - over));
        for (uint32_t i = 0; i < over; ++i) rl[i] = (uint8_t)(hi >> (8u * i));
      }
      ptr -= k;
      ro = (ro - k) & 127u;
    }
    pend = ptr + 64 <= fl;  // the piece [fl - 64, fl) is final
  }
  if (pend) {
    copy64(fl - 64, rl + fro);
    fl -= 64;
    fro = (fro - 64u) & 127u;
  }
  ptr -= 4;
  ro = (ro - 4u) & 127u;  // (the ring offset of ptr - 8 again)
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) rl[(ro + 8u + i) & 127u] = (uint8_t)(x >> (8u * i));  // the state at ptr[0..3]
  while (ptr < fl) {  // the rest, in whole pieces (below ptr: unused scratch of this lane)
    copy64(fl - 64, rl + fro);
    fl -= 64;
    fro = (fro - 64u) & 127u;
  }
  const uint32_t ll = (uint32_t)(end - ptr);
  R.lanelen[(size_t)g * 64 + lane] = ll;
  const uint32_t total = 32u + 2u * npres + 256u + wave_sum(ll);
  if (lane == 0) {
    const uint32_t kind = total >= n ? 0u : 1u;  // raw when rANS is not smaller (src/codec.h:129-135)
    si[0] = kind; si[1] = kind ? total : n; si[2] = npres; si[3] = n;
  }
}
// the literal streams of every block first, then the other four streams (a literal stream is usually many times
// longer: started first, its wavefronts do not trail the kernel's end)
__global__ __launch_bounds__(64) void k_enc_rans(uint32_t blog, TokBufs T, RansBufs R) {
  const uint32_t nbk = gridDim.x / 5u;
  enc_rans_stream(blog, T, R, blockIdx.x % nbk, blockIdx.x / nbk);
}

// ---------------------------------------------------------------- 8. the payload (one workgroup per block)
// out + b*B receives the ZB2 payload (src/codec2.h:140-151) and outlen[b] its size, or 0xFFFFFFFF when the payload
// would not be smaller than the block (the block is then stored).
__device__ inline void wg_copy(uint8_t* d, const uint8_t* s, uint32_t n) {
  typedef uint32_t v4 __attribute__((ext_vector_type(4)));
  const uint32_t full = n & ~15u;
  for (uint32_t i = threadIdx.x * 16; i < full; i += blockDim.x * 16) {
    v4 v;
    __builtin_memcpy(&v, s + i, 16);
    __builtin_memcpy(d + i, &v, 16);
  }
  if (full + threadIdx.x < n) d[full + threadIdx.x] = s[full + threadIdx.x];
}

__device__ __forceinline__ void enc_pack_block(const Block* __restrict__ blocks, uint32_t blog, TokBufs T, RansBufs R,
                                      uint8_t* __restrict__ out, uint32_t* __restrict__ outlen, uint32_t b) {
  __shared__ uint32_t loff[65];
  const Geo G = geo_of(blog);
  const uint32_t tid = threadIdx.x;
  if (blocks[b].raw_len > G.B) { if (tid == 0) outlen[b] = BAD_BLOCK; return; }
  const uint32_t n = blocks[b].raw_len;
  const uint32_t* cn = T.cnt + (size_t)b * 8;
  const uint32_t nob = cn[3];
  uint32_t off[6];
  off[0] = 20;
  for (uint32_t s = 0; s < 5; ++s) off[s + 1] = off[s] + 5 + R.sinfo[((size_t)b * 5 + s) * 4 + 1];
  const uint32_t total = off[5] + nob;
  if (total >= n) { if (tid == 0) outlen[b] = 0xFFFFFFFFu; return; }
  uint8_t* o = out + ((size_t)b << blog);
  if (tid < 20) {
    const uint32_t v = tid < 4 ? cn[0] : tid < 8 ? cn[1] : tid < 12 ? cn[2] : tid < 16 ? nob : cn[4];
    o[tid] = (uint8_t)(v >> (8 * (tid & 3)));
  }
  if (tid == 0) outlen[b] = total;
  for (uint32_t s = 0; s < 5; ++s) {
    const uint32_t g = b * 5 + s;
    const uint32_t* si = R.sinfo + (size_t)g * 4;
    const uint32_t kind = si[0], len = si[1], np = si[2], ns = si[3];
    uint8_t* d = o + off[s];
    if (tid < 5) d[tid] = tid == 0 ? (uint8_t)kind : (uint8_t)(len >> (8 * (tid - 1)));
    const uint8_t* sy = stream_syms(T, G, b, s);
    if (kind == 0) { wg_copy(d + 5, sy, ns); continue; }
    const uint16_t* fq = R.freq + (size_t)g * 256;
    const uint32_t* ll = R.lanelen + (size_t)g * 64;
    // present bitmap, the frequencies of the present symbols in order, the lane lengths
    const uint32_t f = fq[tid];
    const uint64_t pm = ballot(f != 0);
    if ((tid & 7u) == 0) d[5 + (tid >> 3)] = (uint8_t)(pm >> (tid & 63u) & 0xFFu);
    __shared__ uint32_t wpres[4];
    if ((tid & 63u) == 0) wpres[tid >> 6] = popc64(pm);
    __syncthreads();
    uint32_t rank = popc64(pm & lanemask_lt());
    for (uint32_t w = 0; w < (tid >> 6); ++w) rank += wpres[w];
    if (f) { d[5 + 32 + 2 * rank] = (uint8_t)f; d[5 + 33 + 2 * rank] = (uint8_t)(f >> 8); }
    if (tid < 64) {
      const uint32_t l = ll[tid];
      uint8_t* lp = d + 5 + 32 + 2 * np + 4 * tid;
      lp[0] = (uint8_t)l; lp[1] = (uint8_t)(l >> 8); lp[2] = (uint8_t)(l >> 16); lp[3] = (uint8_t)(l >> 24);
      const uint32_t incl = wave_scan_incl(l);
      loff[tid] = incl - l;
    }
    __syncthreads();
    const uint32_t data0 = 5 + 32 + 2 * np + 256;
    const uint32_t LC = lane_cap(G, s);
    const uint8_t* lb = R.lanes + (size_t)b * lanes_per_block(G) + lane_off(G, s);
    // the 64 lanes' bytes side by side, four threads to a lane, 16 bytes each per step (a lane's bytes end at the
    // end of its LC-byte area, being written backwards)
    for (uint32_t j = tid >> 2; j < 64; j += blockDim.x >> 2) {
      const uint32_t l = ll[j];
      const uint8_t* src = lb + (size_t)(j + 1) * LC - l;
      uint8_t* dst = d + data0 + loff[j];
      const uint32_t full = l & ~15u;
      for (uint32_t i = (tid & 3u) * 16u; i < full; i += 64u) {
        uint32_t v[4];
        __builtin_memcpy(v, src + i, 16);
        __builtin_memcpy(dst + i, v, 16);
      }
      for (uint32_t i = full + (tid & 3u); i < l; i += 4u) dst[i] = src[i];
    }
    __syncthreads();
  }
  const uint8_t* ob = (const uint8_t*)(T.offw + (size_t)b * (G.OCAP / 4u + 1u));
  wg_copy(o + off[5], ob, nob);
}
__global__ __launch_bounds__(256) void k_enc_pack(const Block* __restrict__ blocks, uint32_t blog, TokBufs T,
                                                  RansBufs R, uint8_t* __restrict__ out, uint32_t* __restrict__ outlen) {
  enc_pack_block(blocks, blog, T, R, out, outlen, blockIdx.x);
}

}  // namespace zk
