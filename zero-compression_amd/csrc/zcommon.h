// Shared device-side definitions of the MI355X (gfx950) block codec: block geometry, the per-block table that
// every kernel reads, wave64 primitives and byte-level memory helpers.
//
// Geometry follows the ZB2 format's caps (kernels/zero.cl:8-27): a block is B = 2^block_log bytes (16..20),
// at most TCAP = B/4 + 16 tokens (every match is >= 4 bytes), ECAP = B/5 + 64 ext bytes, OCAP = 5B/8 + 8
// offset-bit bytes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZWAVE 64

namespace zk {

struct Geo {
  uint32_t blog, B, TCAP, ECAP, OCAP;
};
__host__ __device__ inline Geo geo_of(uint32_t blog) {
  Geo g;
  g.blog = blog;
  g.B = 1u << blog;
  g.TCAP = g.B / 4u + 16u;
  g.ECAP = g.B / 5u + 64u;
  g.OCAP = 5u * g.B / 8u + 8u;
  return g;
}

// One entry per block of a launch (32 bytes). For decoding: src/comp_len/method locate the block's stored bytes,
// dst/raw_len where its raw bytes go. For encoding: src/raw_len locate the raw bytes, level (1..9) overrides the
// launch's level for this block when not 0; dst is unused.
struct Block {
  uint64_t src;
  uint64_t dst;
  uint32_t raw_len;
  uint32_t comp_len;
  uint32_t method;
  uint32_t level;
};

// a block's length as the encoder kernels take it: raw_len, or 0 when it exceeds the 2^blog slab that the block's
// scratch is sized for (an entry no caller may pass: k_enc_pack then reports the block as BAD_BLOCK and nothing is
// read or written past its slab)
constexpr uint32_t BAD_BLOCK = 0xFFFFFFFEu;
__host__ __device__ inline uint32_t enc_len(const Block& b, uint32_t blog) {
  return b.raw_len <= (1u << blog) ? b.raw_len : 0u;
}

// per-block decoder error bits (err[b]); the host reports any as a damaged block
enum : uint32_t {
  E_HDR = 1u << 0,     // payload shorter than its 20-byte header
  E_COUNTS = 1u << 1,  // header counts impossible for the block
  E_STREAM = 1u << 2,  // stream section cut, wrong kind or wrong raw length
  E_TABLE = 1u << 3,   // rANS frequency table damaged
  E_LANE = 1u << 4,    // rANS lane cut short
  E_TOKEN = 1u << 5,   // a token reads or writes outside its block, or the lengths do not add up
  E_GROUP = 1u << 6,   // a match reads output of its own group (the format forbids it)
  E_METHOD = 1u << 7,  // unknown method, or a stored block whose lengths differ
};

// ---- wave64 primitives
__device__ inline uint32_t lane_id() { return __lane_id(); }
__device__ inline uint64_t ballot(bool p) { return __ballot(p); }
__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }
__device__ inline uint32_t popc64(uint64_t m) { return (uint32_t)__popcll(m); }
// the set bits of m below this lane
__device__ inline uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// v in `lane` (the lane index must be the same in every lane: v_readlane)
__device__ inline uint32_t readlane(uint32_t v, uint32_t lane) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane); }

// DPP forms (gfx9 data-parallel primitives: row shifts inside 16-lane rows, broadcasts of lanes 15 and 31): an
// inclusive scan is six VALU operations with no LDS traffic. Lanes a DPP source does not reach read 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ inline uint32_t dpp0(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, true);
}
__device__ inline uint32_t from_next_lane(uint32_t v) { return dpp0<0x130>(v); }  // wave_shl:1 (lane 63 reads 0)
__device__ inline uint32_t from_prev_lane(uint32_t v) { return dpp0<0x138>(v); }  // wave_shr:1 (lane 0 reads 0)
__device__ inline uint32_t scan_add(uint32_t v) {  // inclusive
  v += dpp0<0x111>(v);       // row_shr:1
  v += dpp0<0x112>(v);       // row_shr:2
  v += dpp0<0x114>(v);       // row_shr:4
  v += dpp0<0x118>(v);       // row_shr:8
  v += dpp0<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp0<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}
__device__ inline uint32_t scan_max(uint32_t v) {  // inclusive (values >= 0: a missing source reads 0)
  v = max(v, dpp0<0x111>(v));
  v = max(v, dpp0<0x112>(v));
  v = max(v, dpp0<0x114>(v));
  v = max(v, dpp0<0x118>(v));
  v = max(v, dpp0<0x142, 0xA>(v));
  v = max(v, dpp0<0x143, 0xC>(v));
  return v;
}
__device__ inline uint32_t max_all(uint32_t v) { return __builtin_amdgcn_readlane(scan_max(v), 63); }

// whole-wavefront scans and reductions, by DPP (every lane of the wavefront must be active)
__device__ inline uint32_t wave_scan_incl(uint32_t v) { return scan_add(v); }
__device__ inline uint32_t wave_scan_excl(uint32_t v) { return scan_add(v) - v; }
__device__ inline uint32_t wave_sum(uint32_t v) { return __builtin_amdgcn_readlane(scan_add(v), 63); }
__device__ inline uint32_t wave_max(uint32_t v) { return max_all(v); }
__device__ inline uint32_t wave_min(uint32_t v) { return ~max_all(~v); }

// ordering of global-memory writes and later reads among the lanes of one wavefront (a compiler barrier; the
// hardware keeps one wavefront's vector memory operations in order)
__device__ inline void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// waits until every vector memory operation of the wavefront has completed (s_waitcnt vmcnt(0); expcnt and lgkmcnt
// at their maximum, so they do not wait). Placed before a loop whose body keeps loads in flight across iterations:
// the compiler's wait at the loop head then no longer has to cover loads issued before the loop, which would make it
// wait for the iteration's own in-flight load as well (the counter completes in issue order).
__device__ inline void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---- byte-level helpers
__device__ inline uint32_t ld32u(const uint8_t* p) {  // 4 bytes at any address, little endian (one unaligned load)
  uint32_t v;
  __builtin_memcpy(&v, p, 4);
  return v;
}

}  // namespace zk
