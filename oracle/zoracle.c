/* The oracle: a CPU restatement of zero's block codec. TEST INFRASTRUCTURE ONLY (the checker): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product never does.
 *
 * Parity pins (what each function is checked against, in tests/):
 *   crc32_update    known answers (crc32("123456789") = 0xCBF43926, crc32("") = 0), Python zlib.crc32 on random data,
 *                   and the compiled reference (oracle/_ref) on golden vectors.
 *   crc32_combine   crc(A||B) == combine(crc(A), crc(B), |B|) against zlib.crc32 of the concatenation.
 *   normalize       invariants (sum 4096, present >= 1) and the compiled reference's frequency tables; tie rules
 *                   are pinned only by the reference vectors (tests/golden/streams.json).
 *   encode_stream   byte-identical to the compiled reference on golden vectors; decode(encode(x)) == x.
 *   decode_stream   decodes reference-written sections; rejects truncated / damaged ones.
 *   cpu_parse + pack_tokens (= cpu_compress_block)
 *                   byte-identical payloads to the compiled reference on the stress corpus at levels 1/5/9.
 *   decompress_block  decodes every reference payload back to its input.
 *   gpu_match / gpu_lazy / gpu_parse
 *                   the GPU encoder's model; no reference run can pin it (the reference GPU is Intel-only and
 *                   nondeterministic), so tests compare it with an independent brute-force restatement on small
 *                   inputs (every candidate scanned directly) and check validity by decoding its payloads with the
 *                   compiled reference decoder.
 */
#include "zoracle.h"

#include <stdlib.h>
#include <string.h>

static uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
static void wr32(uint8_t* p, uint32_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24); }
static void wr16(uint8_t* p, uint32_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
static uint32_t bitlen(uint32_t d) { uint32_t b = 0; while (d >> (b + 1)) ++b; return b; } /* floor(log2 d), d >= 1 */

/* ------------------------------------------------------------------ CRC-32 (src/crc32.h:7-39) */
#define CRC_POLY 0xEDB88320u
static uint32_t crc_tab[256];
static int crc_ready;
static void crc_init(void) {
  if (crc_ready) return;
  for (uint32_t i = 0; i < 256; ++i) {
    uint32_t c = i;
    for (int k = 0; k < 8; ++k) c = (c >> 1) ^ ((c & 1u) ? CRC_POLY : 0u);
    crc_tab[i] = c;
  }
  crc_ready = 1;
}
uint32_t zo_crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
  crc_init();
  uint32_t c = ~crc;
  for (size_t i = 0; i < n; ++i) c = crc_tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
  return ~c;
}
/* a(x) * b(x) mod P in the reflected representation (bit 31 is x^0) */
static uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
  uint32_t r = 0;
  for (uint32_t m = 1u << 31; m; m >>= 1) {
    if (a & m) r ^= b;
    b = (b & 1u) ? (b >> 1) ^ CRC_POLY : b >> 1; /* b *= x */
  }
  return r;
}
/* src/crc32.h:48-65 restated as polynomial arithmetic: crc(A||B) = crc(A) * x^(8|B|) mod P xor crc(B) */
uint32_t zo_crc32_combine(uint32_t a, uint32_t b, uint64_t len_b) {
  if (!len_b) return a;
  uint32_t xp = 1u << 30; /* x^1 */
  uint32_t acc = 1u << 31; /* x^0 */
  uint64_t e = len_b * 8u;
  while (e) {
    if (e & 1u) acc = gf2_mulmod(acc, xp);
    xp = gf2_mulmod(xp, xp);
    e >>= 1;
  }
  return gf2_mulmod(a, acc) ^ b;
}

/* ------------------------------------------------------------------ rANS stream (src/codec.h:41-205) */
#define NL 64u
#define SCALE_BITS 12u
#define SCALE (1u << SCALE_BITS)
#define RANS_L (1u << 23)

void zo_normalize(const uint32_t counts[256], uint32_t total, uint16_t freq[256]) {
  uint32_t sum = 0;
  for (int s = 0; s < 256; ++s) {
    uint32_t f = 0;
    if (counts[s]) {
      f = (uint32_t)(((uint64_t)counts[s] << SCALE_BITS) / total);
      if (!f) f = 1;
    }
    freq[s] = (uint16_t)f;
    sum += f;
  }
  /* too many: one off the first largest (freq > 1) per step */
  while (sum > SCALE) {
    int best = -1;
    for (int s = 0; s < 256; ++s)
      if (freq[s] > 1 && (best < 0 || freq[s] > freq[best])) best = s;
    freq[best]--;
    sum--;
  }
  /* too few: the deficit to the first largest */
  if (sum < SCALE) {
    int best = 0;
    for (int s = 1; s < 256; ++s)
      if (freq[s] > freq[best]) best = s;
    freq[best] = (uint16_t)(freq[best] + (SCALE - sum));
  }
}

size_t zo_encode_stream(const uint8_t* sym, uint32_t n, uint8_t* out) {
  if (n == 0) { out[0] = 0; wr32(out + 1, 0); return 5; }
  uint32_t counts[256] = {0};
  for (uint32_t i = 0; i < n; ++i) counts[sym[i]]++;
  uint16_t freq[256];
  zo_normalize(counts, n, freq);
  uint32_t cum[256], c = 0, present = 0;
  for (int s = 0; s < 256; ++s) { cum[s] = c; c += freq[s]; present += freq[s] != 0; }
  const uint32_t per = (n + NL - 1) / NL;
  const size_t cap = 2u * per + 16u; /* at most 2 bytes per symbol, plus the 4-byte final state */
  uint8_t* tmp = (uint8_t*)malloc(cap * NL);
  uint32_t lane_len[NL];
  size_t lanes_total = 0;
  for (uint32_t j = 0; j < NL; ++j) {
    uint8_t* end = tmp + cap * (j + 1);
    uint8_t* ptr = end;
    uint32_t x = RANS_L;
    if (j < n) {
      /* this lane's symbols j, j+64, ... encoded last to first */
      for (int64_t i = (int64_t)j + (int64_t)((n - 1 - j) / NL) * NL; i >= 0; i -= NL) {
        const uint32_t f = freq[sym[i]];
        const uint32_t xmax = f << (23 - SCALE_BITS + 8);
        while (x >= xmax) { *--ptr = (uint8_t)x; x >>= 8; }
        x = ((x / f) << SCALE_BITS) + (x % f) + cum[sym[i]];
      }
    }
    ptr -= 4;
    wr32(ptr, x);
    lane_len[j] = (uint32_t)(end - ptr);
    lanes_total += lane_len[j];
  }
  const size_t body = 32 + 2 * (size_t)present + 4 * NL + lanes_total;
  if (body >= n) {
    out[0] = 0;
    wr32(out + 1, n);
    memcpy(out + 5, sym, n);
    free(tmp);
    return 5 + (size_t)n;
  }
  out[0] = 1;
  wr32(out + 1, (uint32_t)body);
  uint8_t* o = out + 5;
  memset(o, 0, 32);
  for (int s = 0; s < 256; ++s)
    if (freq[s]) o[s >> 3] |= (uint8_t)(1u << (s & 7));
  o += 32;
  for (int s = 0; s < 256; ++s)
    if (freq[s]) { wr16(o, freq[s]); o += 2; }
  for (uint32_t j = 0; j < NL; ++j) { wr32(o, lane_len[j]); o += 4; }
  for (uint32_t j = 0; j < NL; ++j) {
    memcpy(o, tmp + cap * (j + 1) - lane_len[j], lane_len[j]);
    o += lane_len[j];
  }
  free(tmp);
  return 5 + body;
}

long zo_decode_stream(const uint8_t* p, size_t avail, uint32_t n, uint8_t* sym) {
  if (avail < 5) return -1;
  const uint32_t kind = p[0], len = rd32(p + 1);
  const uint8_t* body = p + 5;
  if (avail - 5 < len) return -2;
  if (kind == 0) {
    if (len != n) return -3;
    memcpy(sym, body, n);
    return 5 + (long)len;
  }
  if (kind != 1) return -4;
  if (len < 32 + 4 * NL) return -5;
  uint32_t np = 0;
  for (int b = 0; b < 32; ++b) np += (uint32_t)__builtin_popcount(body[b]);
  if (len < 32 + 2 * np + 4 * NL) return -6;
  uint32_t freq[256], cum[256], c = 0;
  const uint8_t* fp = body + 32;
  for (int s = 0; s < 256; ++s) {
    freq[s] = 0;
    if ((body[s >> 3] >> (s & 7)) & 1u) { freq[s] = (uint32_t)fp[0] | (uint32_t)fp[1] << 8; fp += 2; }
    cum[s] = c;
    c += freq[s];
  }
  if (c != SCALE) return -7;
  uint8_t slot[SCALE];
  for (int s = 0; s < 256; ++s)
    for (uint32_t k = 0; k < freq[s]; ++k) slot[cum[s] + k] = (uint8_t)s;
  const uint8_t* lens = fp;
  const uint8_t* data = lens + 4 * NL;
  const uint8_t* body_end = body + len;
  for (uint32_t j = 0; j < NL; ++j) {
    const uint32_t ll = rd32(lens + 4 * j);
    if ((size_t)(body_end - data) < ll || ll < 4) return -8;
    const uint8_t* in = data + 4;
    const uint8_t* in_end = data + ll;
    uint32_t x = rd32(data);
    for (uint32_t i = j; i < n; i += NL) {
      const uint32_t sl = x & (SCALE - 1);
      const uint32_t s = slot[sl];
      sym[i] = (uint8_t)s;
      x = freq[s] * (x >> SCALE_BITS) + sl - cum[s];
      while (x < RANS_L) {
        if (in >= in_end) return -9;
        x = (x << 8) | *in++;
      }
    }
    data += ll;
  }
  return 5 + (long)len;
}

/* ------------------------------------------------------------------ CPU parse (src/codec2.h:43-81) */
size_t zo_cpu_parse(const uint8_t* in, uint32_t n, int depth, zo_token* tokens) {
  size_t nt = 0;
  if (n < 4) return 0;
  enum { HB = 16 };
  int32_t* head = (int32_t*)malloc(sizeof(int32_t) << HB);
  int32_t* prev = (int32_t*)malloc(sizeof(int32_t) * n);
  for (uint32_t i = 0; i < (1u << HB); ++i) head[i] = -1;
  uint32_t inserted = 0;
#define H4(i) ((rd32(in + (i)) * 2654435761u) >> (32 - HB))
  /* the longest match at i among `depth` chain entries (the nearest wins a tie); 0 when under 4 bytes */
#define BEST_AT(i, bd_out, res)                                                          \
  do {                                                                                    \
    const uint32_t at_ = (i);                                                             \
    for (; inserted < at_; ++inserted) {                                                  \
      if (inserted + 4 > n) continue;                                                     \
      const uint32_t h_ = H4(inserted);                                                   \
      prev[inserted] = head[h_];                                                          \
      head[h_] = (int32_t)inserted;                                                       \
    }                                                                                     \
    uint32_t best_ = 0;                                                                   \
    if (at_ + 4 <= n) {                                                                   \
      const uint32_t lim_ = (n - at_ < 4096u) ? n - at_ : 4096u;                          \
      int32_t j_ = head[H4(at_)];                                                         \
      for (int k_ = 0; k_ < depth && j_ >= 0; ++k_, j_ = prev[j_]) {                      \
        const uint32_t q_ = best_ < lim_ - 1 ? best_ : lim_ - 1;                          \
        if (best_ && in[(uint32_t)j_ + q_] != in[at_ + q_]) continue;                     \
        uint32_t l_ = 0;                                                                  \
        while (l_ < lim_ && in[(uint32_t)j_ + l_] == in[at_ + l_]) ++l_;                  \
        if (l_ > best_) { best_ = l_; bd_out = at_ - (uint32_t)j_; }                      \
      }                                                                                   \
    }                                                                                     \
    res = best_ >= 4 ? best_ : 0;                                                         \
  } while (0)
  uint32_t i = 0;
  while (i < n) {
    uint32_t d = 0, m;
    BEST_AT(i, d, m);
    if (!m) { ++i; continue; }
    if (i + 1 < n) { /* lazy = 1: one position ahead */
      uint32_t d2 = 0, m2;
      BEST_AT(i + 1, d2, m2);
      if (m2 && (int64_t)8 * m2 - bitlen(d2) - 8 > (int64_t)8 * m - bitlen(d)) { ++i; m = m2; d = d2; }
    }
    tokens[nt].i = i; tokens[nt].m = m; tokens[nt].d = d;
    ++nt;
    i += m;
  }
#undef BEST_AT
#undef H4
  free(head);
  free(prev);
  return nt;
}

/* ------------------------------------------------------------------ ZB2 payload (src/codec2.h:84-153) */
size_t zo_payload_cap(uint32_t n) { return 4 * (size_t)n + 4096; }

size_t zo_pack_tokens(const uint8_t* in, uint32_t n, const zo_token* tk, size_t ntk, uint8_t* out) {
  uint8_t* lit = (uint8_t*)malloc(n + 1);
  uint8_t* tok = (uint8_t*)malloc(ntk + 1);
  uint8_t* sym = (uint8_t*)malloc(ntk + 1);
  uint8_t* grp = (uint8_t*)malloc(ntk + 1);
  uint8_t* ext = (uint8_t*)malloc(5 * ntk + 1);
  uint8_t* bits = (uint8_t*)malloc(3 * ntk + 8);
  uint32_t nl = 0, nt = 0, ne = 0, nb = 0, ng = 0;
  uint64_t acc = 0;
  int accn = 0;
  uint32_t gstart = 0, gt = 0, op = 0;
  uint32_t recent[3] = {1, 4, 8}, last[3] = {1, 4, 8};
  for (size_t t = 0; t < ntk; ++t) {
    const uint32_t i = tk[t].i, m = tk[t].m, d = tk[t].d;
    /* a group closes at 255 tokens, or before a match that reads what its own group writes
       (reading only its own literals is allowed; cut_below = 0 means a match is never cut) */
    int close = gt == 255;
    const uint32_t j = i - d;
    if (!close && gt && j < op && j + (m < d ? m : d) > gstart) close = 1;
    if (close) {
      grp[ng++] = (uint8_t)gt;
      gt = 0;
      gstart = op;
      memcpy(recent, last, sizeof recent);
    }
    const uint32_t L = i - op;
    memcpy(lit + nl, in + op, L);
    nl += L;
    tok[nt] = (uint8_t)(((L < 15 ? L : 15) << 4) | (m - 4 < 15 ? m - 4 : 15));
    if (L >= 15) { const uint32_t x = L - 15; ext[ne] = (uint8_t)x; ext[ne + 1] = (uint8_t)(x >> 8); ext[ne + 2] = (uint8_t)(x >> 16); ne += 3; }
    if (m - 4 >= 15) { const uint32_t x = m - 19; ext[ne] = (uint8_t)x; ext[ne + 1] = (uint8_t)(x >> 8); ne += 2; }
    int rep = -1;
    for (int r = 0; r < 3 && rep < 0; ++r)
      if (recent[r] == d) rep = r;
    if (rep >= 0) {
      sym[nt] = (uint8_t)(250 + rep);
    } else {
      const uint32_t b = bitlen(d);
      sym[nt] = (uint8_t)b;
      acc |= (uint64_t)(d - (1u << b)) << accn;
      accn += (int)b;
      while (accn >= 8) { bits[nb++] = (uint8_t)acc; acc >>= 8; accn -= 8; }
    }
    last[2] = last[1]; last[1] = last[0]; last[0] = d;
    ++gt;
    ++nt;
    op = i + m;
  }
  if (gt) grp[ng++] = (uint8_t)gt;
  memcpy(lit + nl, in + op, n - op);
  nl += n - op;
  if (accn) bits[nb++] = (uint8_t)acc;
  uint8_t* o = out;
  wr32(o, nt); wr32(o + 4, nl); wr32(o + 8, ne); wr32(o + 12, nb); wr32(o + 16, ng);
  o += 20;
  o += zo_encode_stream(lit, nl, o);
  o += zo_encode_stream(tok, nt, o);
  o += zo_encode_stream(ext, ne, o);
  o += zo_encode_stream(sym, nt, o);
  o += zo_encode_stream(grp, ng, o);
  memcpy(o, bits, nb);
  o += nb;
  free(lit); free(tok); free(sym); free(grp); free(ext); free(bits);
  return (size_t)(o - out);
}

size_t zo_cpu_compress_block(const uint8_t* in, uint32_t n, int depth, uint8_t* out, int* stored) {
  zo_token* tk = (zo_token*)malloc(sizeof(zo_token) * (n / 4 + 1));
  const size_t ntk = zo_cpu_parse(in, n, depth, tk);
  const size_t len = zo_pack_tokens(in, n, tk, ntk, out);
  free(tk);
  *stored = len >= n;
  return len;
}

/* ------------------------------------------------------------------ ZB2 decoder (src/codec2.h:155-203) */
int zo_decompress_block(const uint8_t* in, size_t in_len, uint8_t* out, uint32_t n) {
  if (in_len < 20) return -1;
  const uint32_t nt = rd32(in), nl = rd32(in + 4), ne = rd32(in + 8), nb = rd32(in + 12), ng = rd32(in + 16);
  if (nl > n || nt > n || ne > 5 * (uint64_t)nt || ng > nt) return -2;
  uint8_t* lit = (uint8_t*)malloc(nl + 1);
  uint8_t* tok = (uint8_t*)malloc(nt + 1);
  uint8_t* ext = (uint8_t*)malloc(ne + 4);
  uint8_t* sym = (uint8_t*)malloc(nt + 1);
  uint8_t* grp = (uint8_t*)malloc(ng + 1);
  int rc = 0;
  const uint8_t* p = in + 20;
  const uint8_t* end = in + in_len;
  long r;
#define DEC(buf, cnt) do { r = zo_decode_stream(p, (size_t)(end - p), (cnt), (buf)); if (r < 0) { rc = -3; goto done; } p += r; } while (0)
  DEC(lit, nl);
  DEC(tok, nt);
  DEC(ext, ne);
  DEC(sym, nt);
  DEC(grp, ng);
#undef DEC
  if ((size_t)(end - p) < nb) { rc = -4; goto done; }
  {
    const uint8_t* bits = p;
    uint32_t o = 0, li = 0, ei = 0, t = 0;
    uint64_t bp = 0;
    uint32_t recent[3] = {1, 4, 8}, last[3] = {1, 4, 8};
    for (uint32_t g = 0; g < ng; ++g) {
      const uint32_t gn = grp[g];
      if (gn == 0 || t + gn > nt) { rc = -5; goto done; }
      for (uint32_t k = 0; k < gn; ++k, ++t) {
        const uint32_t tv = tok[t], sv = sym[t];
        uint32_t L = tv >> 4, M = (tv & 15u) + 4u;
        if (L == 15) {
          if (ei + 3 > ne) { rc = -6; goto done; }
          L += (uint32_t)ext[ei] | (uint32_t)ext[ei + 1] << 8 | (uint32_t)ext[ei + 2] << 16;
          ei += 3;
        }
        if ((tv & 15u) == 15) {
          if (ei + 2 > ne) { rc = -6; goto done; }
          M += (uint32_t)ext[ei] | (uint32_t)ext[ei + 1] << 8;
          ei += 2;
        }
        uint32_t D;
        if (sv >= 250) {
          if (sv > 252) { rc = -7; goto done; }
          D = recent[sv - 250];
        } else {
          if (sv > 19) { rc = -7; goto done; }
          uint32_t x = 0;
          for (uint32_t q = 0; q < sv; ++q, ++bp) {
            if ((bp >> 3) >= nb) { rc = -8; goto done; }
            x |= (uint32_t)((bits[bp >> 3] >> (bp & 7)) & 1u) << q;
          }
          D = (1u << sv) + x;
        }
        if (li + (uint64_t)L > nl || (uint64_t)o + L + M > n || D == 0 || D > o + L) { rc = -9; goto done; }
        memcpy(out + o, lit + li, L);
        o += L;
        li += L;
        for (uint32_t q = 0; q < M; ++q) out[o + q] = out[o + q - D];
        o += M;
        last[2] = last[1]; last[1] = last[0]; last[0] = D;
      }
      memcpy(recent, last, sizeof recent);
    }
    if (t != nt || o + (nl - li) != n) { rc = -10; goto done; }
    memcpy(out + o, lit + li, nl - li);
  }
done:
  free(lit); free(tok); free(ext); free(sym); free(grp);
  return rc;
}

/* ------------------------------------------------------------------ the GPU encoder's model (kernels/zero.cl:44-317)
 * kernels/zero.cl's algorithm with three departures, each one to make a block's output a function of its bytes:
 *   - every bucket is sorted completely by (tag, position) (zero.cl scatters with atomics, :101, and sorts windows
 *     of 16 entries stepping 12, :118-119, so its order within a bucket depends on timing);
 *   - the hash of a position whose fifth byte lies past the block reads that byte as 0 (zero.cl reads the next
 *     block's byte or stale slack, :55);
 *   - buckets take min(block_log - 3, 15) hash bits (zero.cl: block_log - 3, i.e. 17 at 2^20); the tag is the
 *     12 bits after them, as in zero.cl. */
void zo_gpu_level(int level, uint32_t* ncand, uint32_t* lencap) {
  /* src/gpu_codec.h:66-70 */
  if (level <= 2) { *ncand = 2; *lencap = 16; }
  else if (level <= 6) { *ncand = 4; *lencap = 16; }
  else { *ncand = 8; *lencap = 32; }
}
static uint32_t hbits_of(uint32_t block_log) { return block_log - 3 < 15 ? block_log - 3 : 15; }
/* the 5-byte hash of position p (kernels/zero.cl:49-58); the byte past the block reads as 0 */
static uint32_t hash5(const uint8_t* in, uint32_t n, uint32_t p) {
  const uint32_t v = rd32(in + p);
  const uint32_t w = p + 4 < n ? in[p + 4] : 0u;
  return (v * 2654435761u) ^ (w * 0x9E3779B1u + (w >> 3));
}
static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}
/* common prefix of in[a..] and in[b..], up to maxlen (kernels/zero.cl:147-160) */
static uint32_t prefix(const uint8_t* in, uint32_t a, uint32_t b, uint32_t maxlen) {
  uint32_t l = 0;
  while (l < maxlen && in[a + l] == in[b + l]) ++l;
  return l;
}

void zo_gpu_match(const uint8_t* in, uint32_t n, uint32_t block_log, int level, uint32_t* match) {
  uint32_t ncand, lencap;
  zo_gpu_level(level, &ncand, &lencap);
  memset(match, 0, sizeof(uint32_t) * n);
  if (n < 4) return;
  const uint32_t keybits = hbits_of(block_log) + 12;
  const uint32_t count = n - 3;
  /* every position ordered by (bucket, tag, position): zero.cl's srt with every bucket sorted completely */
  uint64_t* srt = (uint64_t*)malloc(sizeof(uint64_t) * count);
  for (uint32_t p = 0; p < count; ++p) srt[p] = (uint64_t)(hash5(in, n, p) >> (32 - keybits)) << 32 | p;
  qsort(srt, count, sizeof(uint64_t), cmp_u64);
  for (uint32_t s = 0; s < count; ++s) {
    const uint32_t p = (uint32_t)srt[s];
    const uint32_t tag = (uint32_t)(srt[s] >> 32) & 0xFFFu;
    const uint32_t maxlen = n - p < lencap ? n - p : lencap;
    /* the candidates: of the ncand slots just before s, those with the same 12-bit tag (the key's low bits, the
       bucket aside) and an earlier position (kernels/zero.cl:186-199) */
    uint32_t cand[8], nc = 0;
    for (uint32_t k = 1; k <= ncand; ++k) {
      if (s < k) break;
      const uint32_t j = (uint32_t)srt[s - k];
      if (((uint32_t)(srt[s - k] >> 32) & 0xFFFu) == tag && j < p) cand[nc++] = j;
    }
    uint32_t best = 0, bestd = 0;
    if (nc) { /* only then (kernels/zero.cl:209): a run of one byte value first, it costs no offset bits (:219-222) */
      if (p > 0 && rd32(in + p - 1) == rd32(in + p)) {
        best = 4 + prefix(in, p + 3, p + 4, maxlen - 4);
        bestd = 1;
      }
      for (uint32_t m = 0; m < nc && best < maxlen; ++m) {
        const uint32_t j = cand[m];
        if (rd32(in + j) != rd32(in + p)) continue;
        if (best >= 4 && in[j + best] != in[p + best]) continue;
        const uint32_t l = 4 + prefix(in, j + 4, p + 4, maxlen - 4);
        const uint32_t d = p - j;
        if (l > best || (l == best && d < bestd)) { best = l; bestd = d; }
      }
    }
    match[p] = best >= 4 ? (best << 20) | (bestd - 1) : 0u;
  }
  free(srt);
}

void zo_gpu_lazy(const uint32_t* match, uint32_t n, uint32_t* lazy) {
  /* kernels/zero.cl:239-255 */
  for (uint32_t p = 0; p < n; ++p) {
    const uint32_t v = match[p], v2 = p + 1 < n ? match[p + 1] : 0u;
    const int32_t l = (int32_t)(v >> 20), l2 = (int32_t)(v2 >> 20);
    uint32_t r = v;
    if (l && l2) {
      const int32_t c = (int32_t)bitlen((v & 0xFFFFFu) + 1), c2 = (int32_t)bitlen((v2 & 0xFFFFFu) + 1);
      if (8 * l2 - c2 - 8 > 8 * l - c) r = 0;
    }
    lazy[p] = r;
  }
}

size_t zo_gpu_parse(const uint8_t* in, uint32_t n, int level, uint32_t* lazy, zo_token* tokens) {
  uint32_t ncand, lencap;
  zo_gpu_level(level, &ncand, &lencap);
  size_t nt = 0;
  uint32_t p = 0;
  while (p < n) {
    const uint32_t v = lazy[p];
    uint32_t l = v >> 20;
    if (l == lencap) { /* finished only where the chain uses it (kernels/zero.cl:259-271): up to 4095 */
      const uint32_t maxlen = n - p < 4095u ? n - p : 4095u;
      if (maxlen > l) {
        const uint32_t j = p - (v & 0xFFFFFu) - 1u;
        const uint32_t x = l + prefix(in, j + l, p + l, maxlen - l);
        if (x != l) { lazy[p] = (x << 20) | (v & 0xFFFFFu); l = x; }
      }
    }
    if (l) {
      tokens[nt].i = p; tokens[nt].m = l; tokens[nt].d = (v & 0xFFFFFu) + 1u;
      ++nt;
      p += l;
    } else {
      ++p;
    }
  }
  return nt;
}

size_t zo_gpu_compress_block(const uint8_t* in, uint32_t n, uint32_t block_log, int level, uint8_t* out, int* stored) {
  uint32_t* match = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  uint32_t* lazy = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  zo_token* tk = (zo_token*)malloc(sizeof(zo_token) * (n / 4 + 1));
  zo_gpu_match(in, n, block_log, level, match);
  zo_gpu_lazy(match, n, lazy);
  const size_t ntk = zo_gpu_parse(in, n, level, lazy, tk);
  const size_t len = zo_pack_tokens(in, n, tk, ntk, out);
  free(match); free(lazy); free(tk);
  *stored = len >= n;
  return len;
}
