// Host CRC-32 (IEEE 802.3: reflected polynomial 0xEDB88320, init and final xor ~0), for what the container
// checksums on the host: the entry table and whole files (src/crc32.h:22-65; src/archive.h:279-287). Blocks are
// checksummed on the GPU. Slice-by-8 tables; crc(A||B) from crc(A), crc(B), |B| by multiplying with x^(8|B|) mod P.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace zero {
namespace crc {

struct Tables {
  uint32_t t[8][256];
  Tables() {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c >> 1) ^ ((c & 1u) ? 0xEDB88320u : 0u);
      t[0][i] = c;
    }
    for (int k = 1; k < 8; ++k)
      for (uint32_t i = 0; i < 256; ++i) t[k][i] = (t[k - 1][i] >> 8) ^ t[0][t[k - 1][i] & 0xFFu];
  }
};
inline const Tables& tables() {
  static const Tables T;
  return T;
}

inline uint32_t update(uint32_t crc, const void* data, size_t n) {
  const Tables& T = tables();
  const uint8_t* p = (const uint8_t*)data;
  uint32_t c = ~crc;
  for (; n >= 8; n -= 8, p += 8) {
    uint32_t a, b;
    memcpy(&a, p, 4);
    memcpy(&b, p + 4, 4);
    a ^= c;
    c = T.t[7][a & 255u] ^ T.t[6][(a >> 8) & 255u] ^ T.t[5][(a >> 16) & 255u] ^ T.t[4][a >> 24] ^ T.t[3][b & 255u] ^
        T.t[2][(b >> 8) & 255u] ^ T.t[1][(b >> 16) & 255u] ^ T.t[0][b >> 24];
  }
  while (n--) c = T.t[0][(c ^ *p++) & 255u] ^ (c >> 8);
  return ~c;
}
inline uint32_t of(const void* data, size_t n) { return update(0, data, n); }

inline uint32_t mulmod(uint32_t a, uint32_t b) {  // reflected: bit 31 is x^0
  uint32_t r = 0;
  for (int i = 31; i >= 0; --i) {
    if ((a >> i) & 1u) r ^= b;
    b = (b >> 1) ^ ((b & 1u) ? 0xEDB88320u : 0u);
  }
  return r;
}
inline uint32_t combine(uint32_t a, uint32_t b, uint64_t len_b) {
  if (!len_b) return a;
  uint32_t acc = 1u << 31, sq = 1u << 23;  // x^0, x^8
  for (uint64_t e = len_b; e; e >>= 1) {
    if (e & 1u) acc = mulmod(acc, sq);
    sq = mulmod(sq, sq);
  }
  return mulmod(a, acc) ^ b;
}

}  // namespace crc
}  // namespace zero
