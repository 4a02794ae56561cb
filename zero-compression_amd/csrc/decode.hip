// ZB2 block decoder for gfx950: the five entropy-coded streams of every block (one wavefront per stream, the 64
// rANS lanes on the 64 lanes), then the group-parallel LZ replay (one wavefront per block, up to 64 tokens of a
// group side by side).
//
// Format and validity rules: src/codec2.h:155-203 (the sequential decoder whose checks are mirrored here) and
// src/codec.h:150-205 (a stream section). The reference GPU decoder is kernels/zero.cl:640-854.
//
// Safety: every read stays inside its block's compressed bytes and every write inside its block's output and
// scratch, whatever the data; damage sets err[b] bits and never faults.
#include "zcommon.h"

namespace zk {

// exactly n bytes from s to d (16 bytes at a time, then the tail); the ranges must not overlap
__device__ inline void copy_exact(uint8_t* d, const uint8_t* s, uint32_t n) {
  typedef uint32_t v4 __attribute__((ext_vector_type(4)));
  uint32_t i = 0;
  for (; i + 16 <= n; i += 16) {
    v4 v;
    __builtin_memcpy(&v, s + i, 16);
    __builtin_memcpy(d + i, &v, 16);
  }
  if (n - i >= 8) { uint64_t v; __builtin_memcpy(&v, s + i, 8); __builtin_memcpy(d + i, &v, 8); i += 8; }
  if (n - i >= 4) { uint32_t v; __builtin_memcpy(&v, s + i, 4); __builtin_memcpy(d + i, &v, 4); i += 4; }
  for (; i < n; ++i) d[i] = s[i];
}

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

// the M bytes of a match whose first D bytes are at sp (D < M: the pattern repeats). out_m is where the match goes.
// Bytes of this match already written are read back from out_m by the same lane.
__device__ inline void copy_match(uint8_t* out_m, const uint8_t* sp, uint32_t D, uint32_t M) {
  if (D >= M) { copy_exact(out_m, sp, M); return; }
  if (D >= 16) {
    copy_exact(out_m, sp, D);
    for (uint32_t i = D; i < M; i += 16) {
      const uint32_t c = M - i < 16 ? M - i : 16;
      copy_exact(out_m + i, out_m + i - D, c);  // (the source ends at or before i: already written)
    }
    return;
  }
  // D < 16: a 16-byte register holding the pattern from its start, written every S bytes (S a multiple of D)
  uint8_t pat[16];
#pragma unroll
  for (uint32_t k = 0; k < 16; ++k) pat[k] = sp[k % D];
  const uint32_t S = (16 / D) * D;
  uint32_t i = 0;
  for (; i + 16 <= M; i += S) __builtin_memcpy(out_m + i, pat, 16);
  for (uint32_t k = 0; i + k < M; ++k) out_m[i + k] = pat[k];
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

// ---------------------------------------------------------------- the streams: one wavefront per (block, stream)
__global__ __launch_bounds__(64) void k_dec_streams(const Block* __restrict__ blocks, const uint8_t* __restrict__ comp,
                                                    uint32_t blog, uint8_t* __restrict__ lit, uint8_t* __restrict__ tok,
                                                    uint8_t* __restrict__ ext, uint8_t* __restrict__ sym,
                                                    uint8_t* __restrict__ grp, uint32_t* __restrict__ err) {
  const Geo G = geo_of(blog);
  const uint32_t b = blockIdx.x / 5u, s = blockIdx.x % 5u;
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
  __shared__ uint8_t slot[4096];     // slot -> symbol
  __shared__ uint32_t lstart[65];    // lane data offsets in the body

  // frequencies of the present symbols, in symbol order; lane k owns symbols 4k..4k+3
  uint32_t np = lane < 32 ? (uint32_t)__popc(body[lane]) : 0u;
  np = wave_sum(np);
  if (len < 32 + 2 * np + 4 * 64) { if (lane == 0) atomicOr(&err[b], E_TABLE); return; }
  const uint32_t bits4 = (body[lane >> 1] >> ((lane & 1u) * 4u)) & 15u;
  const uint32_t rank0 = wave_scan_excl((uint32_t)__popc(bits4));
  uint32_t f[4], fsum = 0;
  {
    uint32_t r = rank0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      f[k] = 0;
      if ((bits4 >> k) & 1u) { f[k] = (uint32_t)body[32 + 2 * r] | (uint32_t)body[33 + 2 * r] << 8; ++r; }
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
  const uint32_t ll = ld32u(body + lbase + 4 * lane);
  // every lane at least 4 bytes and all inside the body (src/codec.h:185-187); with each length capped at the
  // body size (< 2^25, a payload is smaller than its block) the 32-bit running sum cannot wrap
  if (ballot(ll < 4 || ll > len - data0 || len > (1u << 25))) { if (lane == 0) atomicOr(&err[b], E_LANE); return; }
  const uint32_t lend = wave_scan_incl(ll);
  if (readlane(lend, 63) > len - data0) { if (lane == 0) atomicOr(&err[b], E_LANE); return; }
  lstart[lane] = data0 + lend - ll;
  if (lane == 63) lstart[64] = data0 + lend;
  __syncthreads();
  // slot table: lane j fills slots 64j..64j+63
  {
    const uint32_t i0 = lane * 64;
    uint32_t lo = 0, hi = 255;  // the last symbol whose cum <= i0 (that symbol is present and covers i0)
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if ((fc[mid] >> 16) <= i0) lo = mid; else hi = mid - 1;
    }
    uint32_t sy = lo;
    for (uint32_t k = 0; k < 64; ++k) {
      const uint32_t i = i0 + k;
      while ((fc[sy] >> 16) + (fc[sy] & 0xFFFFu) <= i) ++sy;
      slot[i] = (uint8_t)sy;
    }
  }
  __syncthreads();
  // 64 lanes decode side by side: lane j holds symbols j, j+64, ...
  const uint8_t* in = body + lstart[lane];
  const uint8_t* in_end = body + lstart[lane + 1];
  uint32_t x = ld32u(in);
  in += 4;
  bool bad = false;
  for (uint32_t i = lane; i < n; i += 64) {
    const uint32_t sl = x & 4095u;
    const uint32_t sy = slot[sl];
    const uint32_t e = fc[sy];
    dest[i] = (uint8_t)sy;
    x = (e & 0xFFFFu) * (x >> 12) + sl - (e >> 16);
    while (x < (1u << 23)) {
      if (in >= in_end) { bad = true; break; }
      x = (x << 8) | *in++;
    }
    if (bad) break;
  }
  if (ballot(bad) && lane == 0) atomicOr(&err[b], E_LANE);
}

// ---------------------------------------------------------------- the LZ replay: one wavefront per block
__global__ __launch_bounds__(64) void k_dec_lz(const Block* __restrict__ blocks, const uint8_t* __restrict__ comp,
                                               uint32_t blog, const uint8_t* __restrict__ lit,
                                               const uint8_t* __restrict__ tok, const uint8_t* __restrict__ ext,
                                               const uint8_t* __restrict__ sym, const uint8_t* __restrict__ grp,
                                               uint8_t* out, uint32_t* __restrict__ err) {
  const Geo G = geo_of(blog);
  const uint32_t b = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const Block blk = blocks[b];
  uint8_t* o = out + blk.dst;
  const uint8_t* p = comp + blk.src;
  const uint32_t raw = blk.raw_len;
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
  uint32_t r0 = 1, r1 = 4, r2 = 8, l0 = 1, l1 = 4, l2 = 8;
  uint32_t fail = 0;
  for (uint32_t g = 0; g < ng && !fail; ++g) {
    const uint32_t gn = grpb[g];
    if (gn == 0 || t + gn > nt) { fail = E_TOKEN; break; }
    const uint32_t gstart = op, gfirst = t, tend = t + gn;
    while (t < tend) {
      const uint32_t c = min(64u, tend - t);
      const bool on = lane < c;
      const uint32_t tv = on ? tokb[t + lane] : 0u;
      const uint32_t sv = on ? symb[t + lane] : 0u;
      const uint32_t fl = on && (tv >> 4) == 15u, fm = on && (tv & 15u) == 15u;
      const uint32_t nx = 3u * fl + 2u * fm;
      const uint32_t xq = ei + wave_scan_excl(nx);
      bool bad = on && (uint64_t)xq + nx > ne;
      uint32_t L = tv >> 4, M = on ? (tv & 15u) + 4u : 0u;
      if (!bad) {
        if (fl) L += (uint32_t)extb[xq] | (uint32_t)extb[xq + 1] << 8 | (uint32_t)extb[xq + 2] << 16;
        if (fm) { const uint32_t q = xq + 3u * fl; M += (uint32_t)extb[q] | (uint32_t)extb[q + 1] << 8; }
      }
      const bool rep = sv >= 250u;
      if (on && (rep ? sv > 252u : sv > 19u)) bad = true;
      const uint32_t nb_ = on && !rep && sv <= 19u ? sv : 0u;
      const uint32_t bq = bp + wave_scan_excl(nb_);
      uint32_t D = 0;
      if (on && !bad) {
        if (rep) {
          D = sv == 250u ? r0 : sv == 251u ? r1 : r2;
        } else if ((uint64_t)bq + nb_ > nbits_total) {
          bad = true;
        } else {
          uint32_t w = 0;
          const uint32_t at = bq >> 3;
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k)
            if (at + k < nob) w |= (uint32_t)bits[at + k] << (8 * k);
          D = (1u << nb_) + ((w >> (bq & 7u)) & ((1u << nb_) - 1u));
        }
      }
      const uint32_t tot = L + M;
      const uint32_t opk = op + wave_scan_excl(tot);
      const uint32_t lik = li + wave_scan_excl(L);
      const uint32_t ms = opk + L;
      if (on && !bad && ((uint64_t)lik + L > nl || (uint64_t)opk + tot > raw || D == 0 || D > ms)) bad = true;
      const uint32_t src = ms - D;
      // the group rule: a match reads output from before its group, or its own literals (src/codec2.h:1-9)
      const bool gbad = on && !bad && src < opk && t + lane != gfirst && src + min(M, D) > gstart;
      if (ballot(bad || gbad)) { fail = ballot(bad) ? E_TOKEN : E_GROUP; break; }
      if (on) {
        if (L) copy_exact(o + opk, litb + lik, L);
        if (M) {
          const uint8_t* sp = src >= opk ? litb + lik + (src - opk) : o + src;
          copy_match(o + ms, sp, D, M);
        }
      }
      op += wave_sum(tot);
      li += wave_sum(L);
      ei += wave_sum(nx);
      bp += wave_sum(nb_);
      for (uint32_t q = c > 3u ? c - 3u : 0u; q < c; ++q) {
        const uint32_t v = readlane(D, q);
        l2 = l1; l1 = l0; l0 = v;
      }
      t += c;
    }
    r0 = l0; r1 = l1; r2 = l2;
    wave_fence();  // this group's output is read by the next
  }
  if (!fail && (t != nt || op + (nl - li) != raw)) fail = E_TOKEN;
  if (fail) { if (lane == 0) atomicOr(&err[b], fail); return; }
  wave_copy(o + op, litb + li, nl - li);
}

}  // namespace zk
