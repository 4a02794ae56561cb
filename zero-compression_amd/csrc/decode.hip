// ZB2 block decoder for gfx950: the five entropy-coded streams of every block (one wavefront per stream, the 64
// rANS lanes on the 64 lanes), then the group-parallel LZ replay (one wavefront per block, up to 64 tokens of a
// group side by side).
//
// Format and validity rules: src/codec2.h:155-203 (the sequential decoder whose checks are mirrored here) and
// src/codec.h:150-205 (a stream section). The reference GPU decoder is kernels/zero.cl:640-854.
//
// Safety: every read stays inside its block's compressed bytes and every write inside its block's output and
// scratch, whatever the data; damage sets err[b] bits and never faults.
//
// The bodies dec_stream and dec_lz_block are __forceinline__: the plain kernels and the persistent ones (plan.hip)
// share them, and a shared body compiled as a call would take the call convention's registers (k_dec_lz: 260 VGPRs
// instead of 96).
#include "zcommon.h"

namespace zk {

// a wavefront copies n bytes together (16 bytes per lane per step)
__device__ inline void wave_copy(uint8_t* d, const uint8_t* s, uint32_t n) {
  typedef uint32_t v4 __attribute__((ext_vector_type(4)));
  const uint32_t lane = lane_id();
  const uint32_t full = n & ~15u;
  for (uint32_t i = lane * 16; i < full; i += 64 * 16) {
    v4 v;
    __builtin_memcpy(&v, s + i, 16);
    __builtin_memcpy(d + i, &v, 16);
  }
  if (full + lane < n) d[full + lane] = s[full + lane];
}

struct Sections {
  uint32_t nt, nl, ne, nob, ng;
  uint32_t pos, len;  // the wanted section's offset and body length
  uint32_t bits_pos;  // where the offset bits start
  uint32_t err;
};

// the payload header and the five section headers (src/codec2.h:156-168; counts checked as at :161); pos/len of
// section `want` (0..4)
__device__ inline Sections read_sections(const uint8_t* p, uint32_t clen, uint32_t raw, const Geo& G, uint32_t want) {
  Sections S = {};
  if (raw > G.B) { S.err = E_COUNTS; return S; }  // (a block table entry larger than the block size)
  if (clen < 20) { S.err = E_HDR; return S; }
  S.nt = ld32u(p); S.nl = ld32u(p + 4); S.ne = ld32u(p + 8); S.nob = ld32u(p + 12); S.ng = ld32u(p + 16);
  if (S.nl > raw || S.nt > raw || (uint64_t)S.ne > 5ull * S.nt || S.ng > S.nt || S.nt > G.TCAP || S.ne > G.ECAP) {
    S.err = E_COUNTS;
    return S;
  }
  uint32_t pos = 20;
  for (uint32_t s = 0; s < 5; ++s) {
    if (clen - pos < 5) { S.err = E_STREAM; return S; }
    const uint32_t len = ld32u(p + pos + 1);
    if (len > clen - pos - 5) { S.err = E_STREAM; return S; }
    if (s == want) { S.pos = pos; S.len = len; }
    pos += 5 + len;
  }
  if (S.nob > clen - pos) { S.err = E_STREAM; return S; }
  S.bits_pos = pos;
  return S;
}

// the 16 bytes of a lane at lp + at, requested without a branch: the load starts at min(lp + at, lim - 16) (lim - 16
// is inside the payload, which is at least 20 bytes), and d (0..16) is how far the wanted bytes lie past that start;
// shift16 then drops those d bytes and fills zeros, for the bytes past lim
__device__ inline void load16(const uint8_t* lp, uint32_t at, const uint8_t* lim, uint32_t& v0, uint32_t& v1,
                              uint32_t& v2, uint32_t& v3, uint32_t& d) {
  const uint8_t* q = lp + at;
  const uint8_t* a = q + 16 <= lim ? q : lim - 16;
  d = (uint32_t)min((uint64_t)(q - a), (uint64_t)16);
  uint32_t v[4];
  __builtin_memcpy(v, a, 16);
  v0 = v[0]; v1 = v[1]; v2 = v[2]; v3 = v[3];
}
// a when the mask m is all ones, b when it is zero (a bitwise select: no branch)
__device__ inline uint32_t msel(uint32_t m, uint32_t a, uint32_t b) { return (a & m) | (b & ~m); }
// bytes [d, d + 16) of the 32 bytes w0..w7 (d <= 16; bytes past w7 read as 0), selected by the bits of d / 4
__device__ inline void realign16(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t w4, uint32_t w5,
                                 uint32_t w6, uint32_t w7, uint32_t d, uint32_t& v0, uint32_t& v1, uint32_t& v2,
                                 uint32_t& v3) {
  const uint32_t q = d >> 2, r = d & 3u;
  const uint32_t m1 = 0u - (q & 1u), m2 = 0u - ((q >> 1) & 1u), m4 = 0u - (q >> 2);
  const uint32_t t0 = msel(m1, w1, w0), t1 = msel(m1, w2, w1), t2 = msel(m1, w3, w2), t3 = msel(m1, w4, w3),
                 t4 = msel(m1, w5, w4), t5 = msel(m1, w6, w5), t6 = msel(m1, w7, w6);
  const uint32_t u0 = msel(m2, t2, t0), u1 = msel(m2, t3, t1), u2 = msel(m2, t4, t2), u3 = msel(m2, t5, t3),
                 u4 = msel(m2, t6, t4);
  const uint32_t e0 = msel(m4, w4, u0), e1 = msel(m4, w5, u1), e2 = msel(m4, w6, u2), e3 = msel(m4, w7, u3),
                 e4 = msel(m4, 0u, u4);
  v0 = __builtin_amdgcn_alignbyte(e1, e0, r);
  v1 = __builtin_amdgcn_alignbyte(e2, e1, r);
  v2 = __builtin_amdgcn_alignbyte(e3, e2, r);
  v3 = __builtin_amdgcn_alignbyte(e4, e3, r);
}
__device__ inline void shift16(uint32_t& v0, uint32_t& v1, uint32_t& v2, uint32_t& v3, uint32_t d) {
  realign16(v0, v1, v2, v3, 0u, 0u, 0u, 0u, d, v0, v1, v2, v3);
}

// one rANS decoding step of a lane: returns the symbol of state x (slot table, then its freq | cum << 16). The
// lane's next bytes are w, big-endian (the first byte in the top 8 bits); the renormalisation shifts 0, 1 or 2 of
// them into the state: two below 2^15, one below 2^23, read off the leading zeros of the new state (those of
// xs | 2^11, at most 20, so that the count is 2 for every state below 2^15). ob counts the bits taken.
__device__ inline uint32_t rans_step(const uint8_t* slot, const uint32_t* fc, bool valid, uint64_t& w, uint32_t& x,
                                     uint32_t& ob) {
  const uint32_t sl = x & 4095u;
  const uint32_t sy = slot[sl];
  const uint32_t e = fc[sy];
  const uint32_t xs = (e & 0xFFFFu) * (x >> 12) + sl - (e >> 16);
  const uint32_t lz = (uint32_t)__builtin_clz(xs | 0x800u);  // 0..20
  const uint32_t s8 = valid ? __builtin_elementwise_sub_sat(lz, 1u) & 24u : 0u;
  const uint64_t xw = ((uint64_t)xs << 32 | (uint32_t)(w >> 32)) << s8;
  x = (uint32_t)(xw >> 32);
  w <<= s8;
  ob += s8;
  return sy;
}
// the LDS ring of a lane's bytes in the stream decoder: LRING bytes (4 units of 16) and a copy of the first 8 after
constexpr uint32_t LRING = 64;
// unit u (16 bytes v0..v3) into its ring slot, and slot 0's first 8 bytes also after the ring
__device__ inline void ring_unit(uint8_t* mr, uint32_t u, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3) {
  const uint32_t v[4] = {v0, v1, v2, v3};
  __builtin_memcpy(mr + 16u * (u & 3u), v, 16);
  if ((u & 3u) == 0u) __builtin_memcpy(mr + LRING, v, 8);
}
// 4 rANS steps of a lane, their (at most 8) bytes read from its ring at q, which moves past them; symbol 64 (k0 + k)
// + lane of the 8 steps from d goes to d[64 (k0 + k) + lane]: the 64 lanes of a step store 64 consecutive bytes. In
// the last steps of a stream (TAIL), a lane whose symbol index first + 64 (k0 + k) + lane is past n takes no bytes
// and stores nothing.
template <bool TAIL>
__device__ inline void steps4r(const uint8_t* slot, const uint32_t* fc, uint8_t* d, const uint8_t* mr, uint32_t& q,
                               uint32_t& x, uint32_t k0, uint32_t first, uint32_t n) {
  const uint32_t lane = lane_id();
  uint64_t le;
  __builtin_memcpy(&le, mr + (q & (LRING - 1u)), 8);
  uint64_t w = (uint64_t)__builtin_bswap32((uint32_t)le) << 32 | __builtin_bswap32((uint32_t)(le >> 32));
  uint32_t ob = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const bool valid = !TAIL || first + (k0 + k) * 64 + lane < n;
    const uint32_t sy = rans_step(slot, fc, valid, w, x, ob);
    if (!TAIL || valid) d[(k0 + k) * 64 + lane] = (uint8_t)sy;
  }
  q += ob >> 3;
}

// ---------------------------------------------------------------- the streams: one wavefront per (block, stream)
// stream s of block b
__device__ __forceinline__ void dec_stream(const Block* __restrict__ blocks, const uint8_t* __restrict__ comp, uint32_t blog,
                                  uint8_t* __restrict__ lit, uint8_t* __restrict__ tok, uint8_t* __restrict__ ext,
                                  uint8_t* __restrict__ sym, uint8_t* __restrict__ grp, uint32_t* __restrict__ err,
                                  uint32_t b, uint32_t s) {
  const Geo G = geo_of(blog);
  const uint32_t lane = threadIdx.x;
  const Block blk = blocks[b];
  if (blk.method != 1u) return;
  const uint8_t* p = comp + blk.src;
  const Sections S = read_sections(p, blk.comp_len, blk.raw_len, G, s);
  if (S.err) { if (lane == 0 && s == 0) atomicOr(&err[b], S.err); return; }
  const uint8_t kind = p[S.pos];
  const uint32_t len = S.len;
  const uint8_t* body = p + S.pos + 5;
  uint8_t* dest;
  uint32_t n;
  switch (s) {
    case 0: dest = lit + ((size_t)b << blog); n = S.nl; break;
    case 1: dest = tok + (size_t)b * G.TCAP; n = S.nt; break;
    case 2: dest = ext + (size_t)b * G.ECAP; n = S.ne; break;
    case 3: dest = sym + (size_t)b * G.TCAP; n = S.nt; break;
    default: dest = grp + (size_t)b * G.TCAP; n = S.ng; break;
  }
  if (kind == 0) {
    if (len != n) { if (lane == 0) atomicOr(&err[b], E_STREAM); return; }
    wave_copy(dest, body, n);
    return;
  }
  if (kind != 1 || len < 32 + 4 * 64) { if (lane == 0) atomicOr(&err[b], E_STREAM); return; }

  __shared__ uint32_t fc[256];       // freq | cum << 16
  __shared__ uint32_t slot32[1024];  // slot -> symbol (bytes, filled four at a time)
  uint8_t* const slot = (uint8_t*)slot32;

  // frequencies of the present symbols, in symbol order; lane k owns symbols 4k..4k+3
  uint32_t np = lane < 32 ? (uint32_t)__popc(body[lane]) : 0u;
  np = wave_sum(np);
  if (len < 32 + 2 * np + 4 * 64) { if (lane == 0) atomicOr(&err[b], E_TABLE); return; }
  const uint32_t bits4 = (body[lane >> 1] >> ((lane & 1u) * 4u)) & 15u;
  const uint32_t rank0 = wave_scan_excl((uint32_t)__popc(bits4));
  uint32_t f[4], fsum = 0;
  {  // the lane's (at most 4) u16 frequencies in one 8-byte load (the lane lengths follow the table: in bounds)
    uint64_t fw;
    __builtin_memcpy(&fw, body + 32 + 2 * rank0, 8);
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool on = (bits4 >> k) & 1u;
      f[k] = on ? (uint32_t)(fw >> (16u * r)) & 0xFFFFu : 0u;
      r += on ? 1u : 0u;
      fsum += f[k];
    }
  }
  uint32_t c = wave_scan_excl(fsum);
  const uint32_t total = readlane(c + fsum, 63);
  if (total != 4096u) { if (lane == 0) atomicOr(&err[b], E_TABLE); return; }
#pragma unroll
  for (int k = 0; k < 4; ++k) { fc[lane * 4 + k] = f[k] | c << 16; c += f[k]; }
  // lane lengths: each at least 4 bytes, together inside the body (src/codec.h:185-187)
  const uint32_t lbase = 32 + 2 * np;
  const uint32_t data0 = lbase + 4 * 64;
  uint32_t ll;
  __builtin_memcpy(&ll, body + lbase + 4 * lane, 4);
  // every lane at least 4 bytes and all inside the body (src/codec.h:185-187); with each length capped at the
  // body size (< 2^25, a payload is smaller than its block) the 32-bit running sum cannot wrap
  if (ballot(ll < 4 || ll > len - data0 || len > (1u << 25))) { if (lane == 0) atomicOr(&err[b], E_LANE); return; }
  const uint32_t lend = wave_scan_incl(ll);
  if (readlane(lend, 63) > len - data0) { if (lane == 0) atomicOr(&err[b], E_LANE); return; }
  __syncthreads();
  // slot table: lane j fills slots 64j..64j+63
  {
    const uint32_t i0 = lane * 64;
    uint32_t lo = 0, hi = 255;  // the last symbol whose cum <= i0 (that symbol is present and covers i0)
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if ((fc[mid] >> 16) <= i0) lo = mid; else hi = mid - 1;
    }
    uint32_t sy = lo, e = fc[sy], end = (e >> 16) + (e & 0xFFFFu);  // (end: where sy's slots stop)
    for (uint32_t k = 0; k < 64; k += 4) {  // four slots to a 32-bit store
      uint32_t w = 0;
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        while (end <= i0 + k + u) {
          e = fc[++sy];
          end = (e >> 16) + (e & 0xFFFFu);
        }
        w |= sy << (8u * u);
      }
      slot32[(i0 + k) >> 2] = w;
    }
  }
  __syncthreads();
  // 64 lanes decode side by side: lane j holds symbols j, j+64, ... A step needs 0, 1 or 2 more bytes of the lane
  // (x >= 2^11 after the step: 2^15 <= x < 2^23 takes one, x < 2^15 two), so 4 steps take at most 8 and 8 at most
  // 16. A lane's bytes after its 4-byte state go through a ring of LRING = 64 bytes in LDS, 16 at a time (unit u:
  // bytes [16u, 16u + 16), in ring slot u % 4; the first 8 bytes of slot 0 are repeated after the ring, so that any 8
  // bytes read at once): every 4 steps read their 8 bytes at the lane's position with one LDS read. Steps go 8 at a
  // time. A lane whose ring holds at most 48 bytes past its position requests its next unit as 8 steps start, and
  // writes it into the ring when the next 8 start (a load has 8 steps to arrive): the ring then always holds the 16
  // bytes that 8 steps may take, and a unit is written only over bytes already taken. Each step's 64 symbols leave in
  // one store of 64 consecutive bytes (a lane's unit load is waited for 8 steps later, past the stores issued since,
  // which it does not wait for). Reads stop at the payload's end; a lane that consumes more bytes than its length is
  // damaged.
  __shared__ uint8_t lring[64 * (LRING + 8)];
  uint8_t* const mr = lring + lane * (LRING + 8);
  const uint8_t* lp = body + data0 + lend - ll;  // this lane's bytes
  const uint8_t* ub = lp + 4;                      // its unit 0
  const uint32_t lnl = ll;
  const uint8_t* lim = p + blk.comp_len;
  uint32_t x;
  __builtin_memcpy(&x, lp, 4);  // (ll >= 4: inside the lane)
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {  // units 0..3
    uint32_t v0, v1, v2, v3, d;
    load16(ub, 16u * u, lim, v0, v1, v2, v3, d);
    if (d) shift16(v0, v1, v2, v3, d);
    ring_unit(mr, u, v0, v1, v2, v3);
  }
  uint32_t U = 4, q = 0;  // units in the ring; bytes taken
  bool staged = false;    // unit U is on its way in g0..g3 (moved back by gd bytes)
  uint32_t g0 = 0, g1 = 0, g2 = 0, g3 = 0, gd = 0;
  wait_vm_all();  // (the step loop's head then waits for nothing from before the loop)
  const uint32_t nsteps = (n + 63) >> 6, nfull = n >> 6;  // (every lane has a symbol in the first nfull steps)
  uint32_t b0 = 0;
  for (; b0 + 8 <= nfull; b0 += 8) {
    if (staged) {
      if (gd) shift16(g0, g1, g2, g3, gd);
      ring_unit(mr, U, g0, g1, g2, g3);
      ++U;
    }
    staged = 16u * U - q <= 48u;
    if (staged) load16(ub, 16u * U, lim, g0, g1, g2, g3, gd);
    steps4r<false>(slot, fc, dest + b0 * 64, mr, q, x, 0u, 0u, 0u);
    steps4r<false>(slot, fc, dest + b0 * 64, mr, q, x, 4u, 0u, 0u);
  }
  if (staged) {
    if (gd) shift16(g0, g1, g2, g3, gd);
    ring_unit(mr, U, g0, g1, g2, g3);
    ++U;
  }
  if (b0 < nsteps) {  // the last (at most 8) steps, within the ring
    steps4r<true>(slot, fc, dest + b0 * 64, mr, q, x, 0u, b0 * 64, n);
    steps4r<true>(slot, fc, dest + b0 * 64, mr, q, x, 4u, b0 * 64, n);
  }
  if (ballot(q + 4u > lnl) && lane == 0) atomicOr(&err[b], E_LANE);
}
// the literal streams of every block first, then the other four streams: a literal stream is usually many times
// longer than the others, and started first its wavefronts do not trail the kernel's end
__global__ __launch_bounds__(64) void k_dec_streams(const Block* __restrict__ blocks, const uint8_t* __restrict__ comp,
                                                    uint32_t blog, uint8_t* __restrict__ lit, uint8_t* __restrict__ tok,
                                                    uint8_t* __restrict__ ext, uint8_t* __restrict__ sym,
                                                    uint8_t* __restrict__ grp, uint32_t* __restrict__ err) {
  const uint32_t nb = gridDim.x / 5u;
  dec_stream(blocks, comp, blog, lit, tok, ext, sym, grp, err, blockIdx.x % nb, blockIdx.x / nb);
}

// ---------------------------------------------------------------- the LZ replay: one wavefront per block
// 16 bytes as two little-endian halves
struct U128 {
  uint64_t lo, hi;
};
__device__ inline U128 ld128(const uint8_t* p) { U128 r; __builtin_memcpy(&r, p, 16); return r; }
__device__ inline void st128(uint8_t* p, const U128& v) { __builtin_memcpy(p, &v, 16); }
// bytes [0, off) of a, then the first 16 - off bytes of b (off 0..16)
__device__ inline U128 splice(const U128& a, const U128& b, uint32_t off) {
  const uint32_t s = off * 8u;
  const uint32_t s63 = s & 63u, r63 = (64u - s) & 63u;
  const uint64_t bl = s == 0 ? b.lo : s < 64 ? b.lo << s63 : 0ull;
  const uint64_t bh = s == 0 ? b.hi : s < 64 ? (b.hi << s63) | (b.lo >> r63) : s < 128 ? b.lo << s63 : 0ull;
  const uint64_t ml = s == 0 ? 0ull : s < 64 ? ~0ull >> r63 : ~0ull;
  const uint64_t mh = s <= 64 ? 0ull : s < 128 ? ~0ull >> ((128u - s) & 63u) : ~0ull;
  return U128{(a.lo & ml) | (bl & ~ml), (a.hi & mh) | (bh & ~mh)};
}
// a % d for a < 2^21, d >= 1, by the hardware reciprocal (1 ulp: the quotient is off by at most one, then corrected)
__device__ inline uint32_t mod_small(uint32_t a, uint32_t d) {
  const uint32_t q = (uint32_t)((float)a * __builtin_amdgcn_rcpf((float)d));
  int32_t r = (int32_t)(a - q * d);
  r += r < 0 ? (int32_t)d : 0;
  r -= r >= (int32_t)d ? (int32_t)d : 0;
  return (uint32_t)r;
}
__device__ inline U128 shr128(const U128& v, uint32_t nbytes) {  // toward byte 0, nbytes 0..15
  const uint32_t s = nbytes * 8u, s63 = s & 63u, r63 = (64u - s) & 63u;
  if (s == 0) return v;
  if (s < 64) return U128{(v.lo >> s63) | (v.hi << r63), v.hi >> s63};
  return U128{v.hi >> s63, 0ull};
}
// 16 bytes of a pattern of period D < 16 whose first D bytes are in p, starting at period byte ph
__device__ inline U128 pattern16(U128 e, uint32_t ph, uint32_t D) {
  if ((D & (D - 1u)) == 0u) {  // a period of 1, 2, 4 or 8 bytes (runs, arrays of such values): its first D bytes
                               // times a repeating constant (16 bytes are whole periods: the phase is 0)
    const uint64_t v = D >= 8u ? e.lo
                               : (e.lo & ((1ull << ((8u * D) & 63u)) - 1ull)) *
                                     (D == 1u ? 0x0101010101010101ull : D == 2u ? 0x0001000100010001ull
                                                                                : 0x0000000100000001ull);
    return U128{v, v};
  }
  for (uint32_t len = D; len < 16; len *= 2) e = splice(e, e, len);  // e[j] = period[j % D]
  // out[j] = e[ph + j] for j < 16 - ph, then e[ph + j - D] (one period back), which is byte j - (16 - ph) of e
  // shifted down by 16 - D
  return splice(shr128(e, ph), shr128(e, 16u - D), 16u - ph);
}

// q = a / d, r = a % d for a < 2^21, d >= 1 (quotient by the hardware reciprocal, corrected by one either way)
__device__ inline void divmod_small(uint32_t a, uint32_t d, uint32_t& q, uint32_t& r) {
  uint32_t qq = (uint32_t)((float)a * __builtin_amdgcn_rcpf((float)d));
  int32_t rr = (int32_t)(a - qq * d);
  if (rr < 0) { rr += (int32_t)d; --qq; }
  if (rr >= (int32_t)d) { rr -= (int32_t)d; ++qq; }
  q = qq;
  r = (uint32_t)rr;
}
// stores the first len (0..15) bytes of v at d
__device__ inline void st_partial(uint8_t* d, U128 v, uint32_t len) {
  uint64_t w = v.lo;
  if (len & 8u) { __builtin_memcpy(d, &w, 8); w = v.hi; d += 8; }
  if (len & 4u) { const uint32_t x = (uint32_t)w; __builtin_memcpy(d, &x, 4); w >>= 32; d += 4; }
  if (len & 2u) { const uint16_t x = (uint16_t)w; __builtin_memcpy(d, &x, 2); w >>= 16; d += 2; }
  if (len & 1u) *d = (uint8_t)w;
}

// the fields of token k of a chunk of tokens, as the copy needs them (one 32-byte LDS record: two 16-byte reads)
struct alignas(16) TokF {
  uint32_t opk, L, M, D;  // output start, literal length, match length, offset
  uint32_t lik, cs, grp, pad;  // literal start, first copy chunk, group (from 0 in the chunk of tokens)
};

// Chunks of a token: its literals in 16-byte chunks, then its match in chunks that each read one stretch of
// consecutive source bytes. A match whose offset D is at least its length M reads [ms - D, me - D) (ms, me: the
// match's output range); one with 16 <= D < M repeats the period [ms - D, ms), each repetition cut into
// ceil(D/16) chunks; one with D < 16 is written 16 bytes at a time as the period pattern.
__device__ inline uint32_t chunk_count(uint32_t L, uint32_t M, uint32_t D) {
  const uint32_t nlit = (L + 15u) >> 4;
  if (D < 16u || D >= M) return nlit + ((M + 15u) >> 4);
  uint32_t reps, rem;
  divmod_small(M, D, reps, rem);
  return nlit + reps * ((D + 15u) >> 4) + ((rem + 15u) >> 4);
}

// ---------------------------------------------------------------- the LZ replay through an LDS window
// The last RING bytes of a block's output stay in LDS (byte y at ring[y % RING], the first 16 bytes mirrored after
// the end so that any 16 bytes read at once). The copy chunks of 64 tokens are laid out a batch of 192 at a time:
// each one's 16 source bytes are requested from global memory together when they are literals or output older than
// the window, or read from the window when they are recent output; then the chunks are written into the window pass
// by pass, LDS reads and writes only, and the output goes to global memory in 16-byte stores, one chunk of tokens
// later (or, for a chunk of tokens whose output is longer than the window, a batch later). So a batch costs one round
// trip to memory instead of one per group.
constexpr uint32_t RING = 4096;
constexpr uint32_t RROUNDS = 3;  // copy chunks laid out at once: RROUNDS x 64 (a batch writes at most 3072 bytes)

__device__ inline U128 ring_read16(const uint8_t* ring, uint32_t y) {
  U128 v;
  __builtin_memcpy(&v, ring + (y & (RING - 1)), 16);
  return v;
}
// len (1..16) bytes of v at output position y, with the wrap and the mirror kept up
__device__ inline void ring_write(uint8_t* ring, uint32_t y, const U128& v, uint32_t len, bool on) {
  const uint32_t i = y & (RING - 1);
  if (on) {
    if (len == 16u) __builtin_memcpy(ring + i, &v, 16);
    else st_partial(ring + i, v, len);
  }
  const bool wrap = on && i + len > RING, mirror = on && i < 16u;
  if (wrap) st_partial(ring, shr128(v, RING - i), i + len - RING);  // (at most 15 bytes)
  if (mirror) {
    const uint32_t m = min(len, 16u - i);
    if (m == 16u) __builtin_memcpy(ring + RING + i, &v, 16);
    else st_partial(ring + RING + i, v, m);
  }
}

// one copy chunk: where it writes (y0, len) and where its source bytes are (see chunk_count)
struct CDesc {
  uint32_t y0, len, s0, flags;  // flags: group | kind << 8 | pattern << 10 | phase << 11 | D << 16 (D < 16 only)
  uint32_t k;                   // the token (index in the chunk of tokens)
  uint32_t rend;                // the end of what it reads of this chunk of tokens' output (0: nothing)
};
enum : uint32_t { K_PRE = 0, K_RING = 1, K_SLOW = 2, K_NONE = 3 };

__device__ inline CDesc describe(uint32_t cj, bool live, uint32_t k, const TokF* stk, const uint8_t* o,
                                 const uint8_t* litb, uint32_t raw, uint32_t lo, const uint8_t*& src) {
  CDesc d;
  const uint32_t kk = live ? k : 0u;
  const TokF tk = stk[kk];
  const uint32_t opk = tk.opk, L = tk.L, M = tk.M, D = tk.D, lik = tk.lik, cs = tk.cs, grp = tk.grp;
  const uint32_t w = cj - cs;
  const uint32_t nlit = (L + 15u) >> 4;
  const uint32_t ms = opk + L;
  const bool litc = w < nlit;
  const uint32_t wm = w - nlit;  // (a match chunk: its index in the match)
  // literals, or a match that does not overlap its output (D >= M): 16-byte chunks of [ms - D, ms - D + M)
  uint32_t y0 = litc ? opk + 16u * w : ms + 16u * wm;
  uint32_t cend = litc ? ms : ms + M;
  uint32_t s0 = litc ? y0 : y0 - D;
  bool pat = false;
  uint32_t ph = 0;
  if (live && !litc && D < M) {  // a match that overlaps its output: the period pattern, or repetitions
    if (D < 16u) {
      pat = true;
      s0 = ms - D;
      ph = mod_small(16u * wm, D);
    } else {
      uint32_t rep, within;
      divmod_small(wm, (D + 15u) >> 4, rep, within);
      const uint32_t base = ms + rep * D;
      y0 = base + 16u * within;
      cend = min(cend, base + D);
      s0 = y0 - (rep + 1u) * D;
    }
  }
  const uint32_t len = live ? min(16u, cend - y0) : 0u;
  const uint32_t need = pat ? D : len;
  uint32_t kind;
  if (!live) kind = K_NONE;
  else if (s0 >= opk) kind = K_PRE;                                    // its own literals
  else if (s0 + need > opk) kind = K_SLOW;                             // part output, part literals
  else if (s0 >= lo) kind = K_RING;                                    // recent output
  else if (s0 + need > lo || s0 + 16u > raw) kind = K_SLOW;            // across the window's edge, or the block's end
  else kind = K_PRE;                                                   // older output, in global memory
  src = s0 >= opk ? litb + lik + (s0 - opk) : o + s0;
  d.y0 = y0;
  d.len = len;
  d.s0 = s0;
  d.flags = grp | kind << 8 | (pat ? 1u : 0u) << 10 | ph << 11 | (pat ? D : 0u) << 16;
  d.k = kk;
  d.rend = kind == K_PRE || kind == K_NONE ? 0u : min(s0 + need, opk);
  return d;
}

// a copy chunk byte by byte: its source bytes from the literals, the window or global memory
__device__ inline void copy_slow(const CDesc& d, uint8_t* ring, const uint8_t* o, const uint8_t* litb, uint32_t lo,
                                 const TokF* stk) {
  const uint32_t opk = stk[d.k].opk, lik = stk[d.k].lik;
  const bool pat = (d.flags >> 10) & 1u;
  const uint32_t ph = (d.flags >> 11) & 15u, D = d.flags >> 16;
  for (uint32_t i = 0; i < d.len; ++i) {
    const uint32_t si = pat ? d.s0 + mod_small(ph + i, D) : d.s0 + i;
    const uint8_t v8 = si >= opk ? litb[lik + (si - opk)] : si >= lo ? ring[si & (RING - 1)] : o[si];
    const uint32_t yi = (d.y0 + i) & (RING - 1);
    ring[yi] = v8;
    if (yi < 16u) ring[RING + yi] = v8;
  }
}

// output [a, e) of the block, now in the window, to global memory
__device__ inline void ring_flush(const uint8_t* ring, uint8_t* o, uint32_t a, uint32_t e) {
  const uint32_t lane = lane_id();
  const uint32_t A = min((a + 15u) & ~15u, e), E = max(A, e & ~15u);
  if (a + lane < A) o[a + lane] = ring[(a + lane) & (RING - 1)];
  for (uint32_t y = A + 16u * lane; y < E; y += 64u * 16u) st128(o + y, ring_read16(ring, y));
  if (E + lane < e) o[E + lane] = ring[(E + lane) & (RING - 1)];
}

// The chunk of tokens' copy chunks [0, C), which write its output [op0, op_end), through the window. A batch that
// writes [ya, yb) (yb <= ya + 3072) reads the window from lo = yb - RING on (a write at y lands on y - RING, before
// lo), and output before lo comes from global memory: what of it is not there yet, [fl_a, lo), is flushed before the
// batch is laid out. fl_a: the first output position not yet in global memory (updated).
__device__ inline void ring_copy(uint8_t* ring, uint8_t* o, const uint8_t* litb, uint32_t raw, uint32_t op0,
                                 uint32_t op_end, uint32_t C, bool on, uint32_t ccs_lane, const TokF* stk,
                                 uint32_t* sown, uint32_t& fl_a) {
  const uint32_t lane = lane_id();
  uint32_t ya = op0;  // where the batch's output starts
  for (uint32_t base = 0; base < C; base += 64 * RROUNDS) {
    const uint32_t yb = min(op_end, ya + 16u * 64u * RROUNDS);  // (at least where it ends)
    const uint32_t lo = yb > RING ? yb - RING : 0u;
    if (lo > fl_a) {  // (only in a chunk of tokens whose output is longer than the window)
      ring_flush(ring, o, fl_a, lo);
      fl_a = lo;
      wave_fence();
    }
    CDesc d[RROUNDS];
    U128 pre[RROUNDS];
#pragma unroll
    for (uint32_t r = 0; r < RROUNDS; ++r) {
      const uint32_t c0 = base + 64 * r;
      if (c0 >= C) {  // (uniform: no chunks left for this round)
        d[r] = CDesc{0u, 0u, 0u, K_NONE << 8, 0u, 0u};
        pre[r] = U128{0, 0};
        continue;
      }
      // the token of chunk c0 + lane: the tokens that start at or before c0, plus those that start in slots 1..lane
      sown[lane] = 0;
      wave_fence();
      const uint32_t jk = ccs_lane - c0;  // (wraps for a token that starts before c0)
      if (on && jk < 64u) sown[jk] = 1u;
      wave_fence();
      const uint64_t starts = ballot(sown[lane] != 0u);
      wave_fence();
      const uint32_t k = popc64(ballot(on && ccs_lane <= c0)) - 1u + mbcnt64(starts >> 1);
      const uint32_t cj = c0 + lane;
      const uint8_t* src;
      d[r] = describe(cj, cj < C, k, stk, o, litb, raw, lo, src);
      pre[r] = ld128(((d[r].flags >> 8) & 3u) == K_PRE ? src : litb);
    }
    // Round by round, and inside a round pass by pass: every chunk whose source bytes are all written goes, that is
    // one that reads output before the first chunk not yet written (chunks are in output order, so every byte
    // before that chunk is in the window), or its own literals, or older output in global memory. The group rule
    // makes the first chunk not yet written always ready (it reads earlier groups' output or its own literals), so
    // a round takes at most 64 passes, and usually about two: a pass is bounded by the dependences of the data,
    // not by its groups. (The branches below are plain divergent ifs: the compiler skips a region that no lane
    // enters, where an explicit ballot would cost two vector instructions.)
#pragma unroll
    for (uint32_t r = 0; r < RROUNDS; ++r) {
      const uint32_t kind = (d[r].flags >> 8) & 3u;
      bool done = kind == K_NONE;
      for (uint32_t it = 0; it < 64u; ++it) {
        const uint64_t left = ballot(!done);
        if (!left) break;
        const uint32_t F = readlane(d[r].y0, (uint32_t)__builtin_ctzll(left));
        const bool act = !done && d[r].rend <= F;
        done = done || act;
        U128 v = pre[r];
        if (act && kind == K_RING) v = ring_read16(ring, d[r].s0);
        const bool pat = (d[r].flags >> 10) & 1u;
        if (act && pat && kind != K_SLOW) v = pattern16(v, (d[r].flags >> 11) & 15u, d[r].flags >> 16);
        ring_write(ring, d[r].y0, v, d[r].len, act && kind != K_SLOW);
        if (act && kind == K_SLOW) copy_slow(d[r], ring, o, litb, lo, stk);
        wave_fence();  // (this pass's bytes are in the window before the next pass reads them)
      }
    }
    // (every load of this batch has been used; saying so keeps the compiler from waiting at the loop's head, which
    // would also wait for the stores and loads issued before the loop)
    wait_vm_all();
    ya = readlane(d[RROUNDS - 1].y0 + d[RROUNDS - 1].len, 63);  // (used only when a full batch follows)
  }
}

// Tokens go 64 at a time: a chunk may hold the end of one group and several whole groups. The fields, positions
// and offsets of the whole chunk are computed together (scans by DPP), with the next chunk's bytes already
// requested; then its copy chunks go through the window (ring_copy: a match reads output from before its group only,
// so the tokens of one group copy side by side, and in practice much more goes at once). Checks follow the
// sequential decoder (src/codec2.h:173-202), plus the group rule.
__device__ __forceinline__ void dec_lz_block(const Block* __restrict__ blocks, const uint8_t* __restrict__ comp, uint32_t blog,
                                    const uint8_t* __restrict__ lit, const uint8_t* __restrict__ tok,
                                    const uint8_t* __restrict__ ext, const uint8_t* __restrict__ sym,
                                    const uint8_t* __restrict__ grp, uint8_t* out, uint32_t* __restrict__ err,
                                    uint32_t b) {
  __shared__ uint32_t sflag[64];  // token k of the chunk starts a group
  __shared__ uint32_t sgs[64];    // output start of the chunk's k-th group
  __shared__ uint32_t sgk[64];    // first token (in the chunk) of the chunk's k-th group
  __shared__ TokF stk[64];          // the chunk of tokens' fields
  __shared__ uint32_t sown[64];
  __shared__ uint8_t ring[RING + 16];  // the block's most recent output
  const Geo G = geo_of(blog);
  const uint32_t lane = threadIdx.x;
  const Block blk = blocks[b];
  uint8_t* o = out + blk.dst;
  const uint8_t* p = comp + blk.src;
  const uint32_t raw = blk.raw_len;
  if (raw > G.B) { if (lane == 0) atomicOr(&err[b], E_COUNTS); return; }  // (before any access: see read_sections)
  if (blk.method == 0u) {
    if (blk.comp_len != raw) { if (lane == 0) atomicOr(&err[b], E_METHOD); return; }
    wave_copy(o, p, raw);
    return;
  }
  if (blk.method != 1u) { if (lane == 0) atomicOr(&err[b], E_METHOD); return; }
  if (err[b]) return;  // (the streams kernel found damage)
  const Sections S = read_sections(p, blk.comp_len, raw, G, 0);
  if (S.err) return;
  const uint8_t* litb = lit + ((size_t)b << blog);
  const uint8_t* tokb = tok + (size_t)b * G.TCAP;
  const uint8_t* extb = ext + (size_t)b * G.ECAP;
  const uint8_t* symb = sym + (size_t)b * G.TCAP;
  const uint8_t* grpb = grp + (size_t)b * G.TCAP;
  const uint8_t* bits = p + S.bits_pos;
  const uint32_t nt = S.nt, nl = S.nl, ne = S.ne, nob = S.nob, ng = S.ng;
  const uint64_t nbits_total = (uint64_t)nob * 8u;

  uint32_t t = 0, op = 0, li = 0, ei = 0, bp = 0;
  uint32_t g = 0, grem = 0, gs_open = 0;  // groups started so far; tokens left in the open group; its output start
  uint32_t r0 = 1, r1 = 4, r2 = 8, l0 = 1, l1 = 4, l2 = 8;
  uint32_t fl_a = 0, fl_e = 0;  // output [fl_a, fl_e) is in the window and not yet in global memory
  uint32_t fail = 0;
  uint32_t tv_n = lane < nt ? tokb[lane] : 0u, sv_n = lane < nt ? symb[lane] : 0u;
  uint32_t gw_n = lane < ng ? grpb[lane] : 0u;
  while (t < nt) {
    const uint32_t c = min(64u, nt - t);
    const bool on = lane < c;
    const uint32_t tv = on ? tv_n : 0u, sv = on ? sv_n : 0u;
    // the groups: lane j holds group g+j's size and its end, counted from the chunk's first token
    const bool gvalid = g + lane < ng;
    const uint32_t gsz = gvalid ? gw_n : 0u;
    const uint32_t gend = grem + scan_add(gsz);
    const uint32_t gbeg = gend - gsz;
    const bool starts = gvalid && gbeg < c;
    const uint32_t cover = max_all(gvalid ? gend : grem);
    if (ballot(starts && gsz == 0) || cover < c) { fail = E_TOKEN; break; }
    sflag[lane] = 0;
    wave_fence();
    if (starts) sflag[gbeg] = 1;
    wave_fence();  // (LDS handoffs between lanes: the wavefront runs them in order, the fence keeps the compiler's)
    const uint64_t startm = ballot(starts);
    const uint32_t nstart = popc64(startm);
    const uint32_t first = on ? sflag[lane] : 0u;
    const uint32_t lastend = nstart ? readlane(gend, nstart - 1) : grem;
    const uint32_t grem2 = lastend - c;
    const uint32_t g2 = g + nstart;
    // The previous chunk's output leaves the window here, at the start of this chunk's parse rather than at the end
    // of its own copy, and the next chunk's token bytes are requested (clamped indexes, no branch around the loads;
    // lanes past nt or ng are masked at the loop's head: lane >= c, g + lane >= ng). They complete together with this
    // chunk's ext and offset-bit loads below, in one memory round trip.
    ring_flush(ring, o, fl_a, fl_e);
    fl_a = fl_e = 0;
    wave_fence();
    {
      const uint32_t tn = min(t + c + lane, nt - 1u), gn = ng ? min(g2 + lane, ng - 1u) : 0u;
      tv_n = tokb[tn];
      sv_n = symb[tn];
      gw_n = grpb[gn];
    }
    // lengths: L >= 15 has L-15 in 3 ext bytes, then M-4 >= 15 has M-19 in 2 (src/codec2.h:178-180); offsets: b
    // raw bits for symbol b (0..19), or one of the group before's last three (250..252). One scan carries the
    // token's ext bytes (bits 0-9), offset bits (10-20) and group start flag (21-27).
    const uint32_t fl = on && (tv >> 4) == 15u, fm = on && (tv & 15u) == 15u;
    const uint32_t nx = 3u * fl + 2u * fm;
    const bool rep = sv >= 250u;
    bool bad = on && (rep ? sv > 252u : sv > 19u);
    const uint32_t nb_ = on && !rep && sv <= 19u ? sv : 0u;
    const uint32_t pk = scan_add(nx | nb_ << 10 | first << 21);
    const uint32_t xq = ei + ((pk & 1023u) - nx);
    const uint32_t bq = bp + (((pk >> 10) & 2047u) - nb_);
    const uint32_t rg = (pk >> 21) - (grem == 0 ? 1u : 0u);  // the token's group, from 0 in the chunk
    const uint32_t tail = readlane(pk, 63);
    if (on && (uint64_t)xq + nx > ne) bad = true;
    uint32_t L = tv >> 4, M = on ? (tv & 15u) + 4u : 0u;
    // the token's ext bytes (at most 5) and offset bits (at most 19), requested together by every lane at clamped
    // places (no branch around the loads: one memory round trip for both), used below where they count. The ext
    // load may read 8 bytes from ne (<= ECAP): the decoder scratch goes on after a block's ext bytes.
    uint64_t xw;
    __builtin_memcpy(&xw, extb + min(xq, ne), 8);
    uint32_t bw = 0;
    {
      const uint32_t at = bq >> 3;
      if (nob >= 4) {  // 4 bytes at a clamped address, shifted: bytes past the section read as 0
        const uint32_t a = min(at, nob - 4u);
        uint32_t v;
        __builtin_memcpy(&v, bits + a, 4);
        bw = at - a < 4u ? v >> (8u * (at - a)) : 0u;
      } else {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
          if (at + k < nob) bw |= (uint32_t)bits[at + k] << (8 * k);
      }
    }
    if (!bad && (fl | fm)) {
      if (fl) L += (uint32_t)xw & 0xFFFFFFu;
      if (fm) M += (uint32_t)(xw >> (24u * fl)) & 0xFFFFu;
    }
    uint32_t D = 0;
    if (on && !bad && !rep) {
      if ((uint64_t)bq + nb_ > nbits_total) bad = true;
      else D = (1u << nb_) + ((bw >> (bq & 7u)) & ((1u << nb_) - 1u));
    }
    // repeated offsets: the recent offsets of a group are those of the three tokens before its first one. So a
    // repeat j (symbol 250 + j) in a group that starts in this chunk at token gsk takes the offset of token
    // gsk - 1 - j: one of this chunk (resolved first when it is a repeat itself), or one of the last three before
    // the chunk (l0..l2); in the group still open from the chunk before it takes r0..r2, that group's recent offsets
    if (first) sgk[rg] = lane;
    wave_fence();
    const bool cont = rg == 0 && grem > 0;  // (a token of the group open from the chunk before)
    const uint32_t gsk = cont ? 0u : sgk[rg & 63u];
    if (ballot(on && rep && !bad)) {
      const bool rp = on && rep && !bad;
      const uint32_t j = rp ? sv - 250u : 0u;
      const uint32_t ix = j - gsk;  // (for a source before the chunk: 0..2)
      bool pend = rp && !cont && gsk > j;
      if (rp) D = cont ? (j == 0 ? r0 : j == 1 ? r1 : r2) : (ix == 0 ? l0 : ix == 1 ? l1 : l2);
      const uint32_t from = pend ? gsk - 1u - j : lane;
      while (ballot(pend)) {  // one link of a chain of repeats per pass (each source is an earlier token)
        const uint32_t v = (uint32_t)__shfl((int)D, (int)from, 64);
        const uint32_t sp = (uint32_t)__shfl(pend ? 1 : 0, (int)from, 64);
        if (pend && !sp) { D = v; pend = false; }
      }
    }
    const uint32_t tot = L + M;
    const uint32_t itot = scan_add(tot), il = scan_add(L);
    const uint32_t opk = op + itot - tot;
    const uint32_t lik = li + il - L;
    const uint32_t ms = opk + L;
    if (first) sgs[rg] = opk;
    wave_fence();
    const uint32_t gstart = (rg == 0 && grem > 0) ? gs_open : sgs[rg & 63u];
    if (on && !bad && ((uint64_t)lik + L > nl || (uint64_t)opk + tot > raw || D == 0 || D > ms)) bad = true;
    const uint32_t src = ms - D;
    // the group rule: a match reads output from before its group, or its own literals (src/codec2.h:1-9)
    const bool gbad = on && !bad && src < opk && !first && src + min(M, D) > gstart;
    if (ballot(bad || gbad)) { fail = ballot(bad) ? E_TOKEN : E_GROUP; break; }
    const uint32_t cnt = on ? chunk_count(L, M, D) : 0u;
    const uint32_t cinc = scan_add(cnt), ccs = cinc - cnt;
    stk[lane] = TokF{opk, L, M, D, lik, ccs, rg, 0u};
    wave_fence();
    if (grem2 && nstart) gs_open = readlane(opk, readlane(gbeg, nstart - 1));  // (the open group's first token)
    {  // recent offsets carried on: those of the group that stays open (when it started here), then l0..l2
      const uint32_t gl = readlane(gsk, c - 1);
      if (!readlane(cont ? 1u : 0u, c - 1)) {
        const uint32_t a0 = gl >= 1u ? readlane(D, gl - 1u) : l0;
        const uint32_t a1 = gl >= 2u ? readlane(D, gl - 2u) : gl == 1u ? l0 : l1;
        const uint32_t a2 = gl >= 3u ? readlane(D, gl - 3u) : gl == 2u ? l0 : gl == 1u ? l1 : l2;
        r0 = a0; r1 = a1; r2 = a2;
      }
      const uint32_t b0 = readlane(D, c - 1u);
      const uint32_t b1 = c >= 2u ? readlane(D, c - 2u) : l0;
      const uint32_t b2 = c >= 3u ? readlane(D, c - 3u) : c == 2u ? l0 : l1;
      l0 = b0; l1 = b1; l2 = b2;
    }
    const uint32_t op_end = op + readlane(itot, 63);
    op = op_end;
    li += readlane(il, 63);
    ei += tail & 1023u;
    bp += (tail >> 10) & 2047u;
    t += c;
    g = g2;
    grem = grem2;
    const uint32_t op0 = op_end - readlane(itot, 63);  // (the chunk's output: [op0, op_end))
    fl_a = op0;
    ring_copy(ring, o, litb, raw, op0, op_end, readlane(cinc, 63), on, ccs, stk, sown, fl_a);
    fl_e = op_end;  // ([fl_a, fl_e) to global memory with the next chunk, or after the loop)
    wave_fence();
  }
  if (!fail && (g != ng || grem != 0 || op + (nl - li) != raw)) fail = E_TOKEN;
  if (fail) { if (lane == 0) atomicOr(&err[b], fail); return; }
  wave_fence();
  ring_flush(ring, o, fl_a, fl_e);
  wave_copy(o + op, litb + li, nl - li);
}
// at most 96 VGPRs, so that 5 wavefronts fit a SIMD (LDS allows 22 per CU): left to itself the compiler takes 97,
// which leaves 4, and 16384 blocks of 2^18 then replay 7 % slower (5.26 against 4.90 ms)
__global__ __launch_bounds__(64, 5) void k_dec_lz(const Block* __restrict__ blocks, const uint8_t* __restrict__ comp,
                                               uint32_t blog, const uint8_t* __restrict__ lit,
                                               const uint8_t* __restrict__ tok, const uint8_t* __restrict__ ext,
                                               const uint8_t* __restrict__ sym, const uint8_t* __restrict__ grp,
                                               uint8_t* out, uint32_t* __restrict__ err) {
  dec_lz_block(blocks, comp, blog, lit, tok, ext, sym, grp, out, err, blockIdx.x);
}

}  // namespace zk
