// Deterministic synthetic inputs for the codec's workloads (host code, no GPU). Exported through a C ABI in
// libzerogen.so so bench.py and the tests produce the same bytes from the same seed.
//
// PRNG: xoshiro256** seeded through SplitMix64. U = ((next >> 11) + 1) * 2^-53 lies in (0, 1].
// geometric(mean mu) = 1 + floor(ln U / ln(1 - 1/mu)) on {1, 2, ...}; Zipf(s, 256) by inverse CDF, byte = rank - 1.
//
//   zgen_lz         LZ-friendly stream: with p = 0.5 a literal run (length geometric(16), Zipf(1.1) bytes) or a
//                   back-copy (length min(4 + geometric(20), 258), offset min(1 + geometric(4096), 2^18, written));
//                   a literal run when nothing is written yet. Overlapping copies repeat their pattern.
//   zgen_random     uniform bytes (the little-endian bytes of successive next() values).
//   zgen_exec       4 KiB pages: 30 % zero, 40 % little-endian uint32 arrays v[i] = v[i-1] + round(N(0, sigma = 64))
//                   from a uniform start, 30 % filled with copies (length uniform 16..512) of earlier data starting at
//                   page-aligned offsets of earlier pages (random bytes when there is no earlier page).
//   zgen_lognormal_sizes  clip(round(exp(ln(median) + sigma * Z)), lo, hi), Z standard normal (Box-Muller).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double u01() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }  // (0, 1]
  uint64_t geometric(double mean) { return 1 + (uint64_t)std::floor(std::log(u01()) / std::log(1.0 - 1.0 / mean)); }
  double normal() {  // Box-Muller, one value per call
    const double a = u01(), b = u01();
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
  }
};

struct Zipf {
  double cdf[256];
  uint16_t guide[4096];  // the first rank whose CDF exceeds k/4096: the inverse CDF starts its scan there
  explicit Zipf(double s) {
    double t = 0;
    for (int r = 1; r <= 256; ++r) { t += std::pow((double)r, -s); cdf[r - 1] = t; }
    for (double& c : cdf) c /= t;
    cdf[255] = 2.0;  // (u < 1 always lands)
    int r = 0;
    for (int k = 0; k < 4096; ++k) {
      while (cdf[r] <= k / 4096.0) ++r;
      guide[k] = (uint16_t)r;
    }
  }
  uint8_t draw(Rng& g) const {
    const double u = g.u01() - 1e-300;  // [0, 1)
    int r = guide[(int)(u * 4096.0)];
    while (cdf[r] <= u) ++r;  // the first rank whose CDF exceeds u
    return (uint8_t)r;
  }
};
}  // namespace

extern "C" {

void zgen_lz(uint64_t seed, uint8_t* out, uint64_t n) {
  Rng g(seed);
  static const Zipf zipf(1.1);
  uint64_t w = 0;
  while (w < n) {
    const bool literal = w == 0 || g.u01() <= 0.5;
    if (literal) {
      uint64_t len = g.geometric(16.0);
      if (len > n - w) len = n - w;
      for (uint64_t i = 0; i < len; ++i) out[w + i] = zipf.draw(g);
      w += len;
    } else {
      uint64_t len = 4 + g.geometric(20.0);
      if (len > 258) len = 258;
      uint64_t off = 1 + g.geometric(4096.0);
      if (off > (1u << 18)) off = 1u << 18;
      if (off > w) off = w;
      if (len > n - w) len = n - w;
      for (uint64_t i = 0; i < len; ++i) out[w + i] = out[w + i - off];
      w += len;
    }
  }
}

void zgen_random(uint64_t seed, uint8_t* out, uint64_t n) {
  Rng g(seed);
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const uint64_t v = g.next();
    memcpy(out + i, &v, 8);
  }
  if (i < n) {
    const uint64_t v = g.next();
    memcpy(out + i, &v, n - i);
  }
}

void zgen_exec(uint64_t seed, uint8_t* out, uint64_t n) {
  Rng g(seed);
  const uint64_t P = 4096;
  for (uint64_t page = 0; page * P < n; ++page) {
    const uint64_t base = page * P;
    const uint64_t len = n - base < P ? n - base : P;
    uint8_t* o = out + base;
    const double kind = g.u01();
    if (kind <= 0.3) {
      memset(o, 0, len);
    } else if (kind <= 0.7) {
      uint32_t v = (uint32_t)(g.next() >> 32);
      for (uint64_t i = 0; i < len; i += 4) {
        v += (uint32_t)(int32_t)std::lround(64.0 * g.normal());
        uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        memcpy(o + i, b, len - i < 4 ? len - i : 4);
      }
    } else if (page == 0) {
      for (uint64_t i = 0; i < len; ++i) o[i] = (uint8_t)g.next();
    } else {
      uint64_t i = 0;
      while (i < len) {
        uint64_t c = 16 + g.next() % 497;  // 16..512
        const uint64_t src = (g.next() % page) * P;
        if (c > len - i) c = len - i;
        for (uint64_t k = 0; k < c; ++k) o[i + k] = out[src + k];  // (src + c <= base + i: earlier data)
        i += c;
      }
    }
  }
}

void zgen_lognormal_sizes(uint64_t seed, double median, double sigma, uint64_t lo, uint64_t hi, uint64_t* sizes,
                          uint64_t count) {
  Rng g(seed);
  const double mu = std::log(median);
  for (uint64_t i = 0; i < count; ++i) {
    double v = std::exp(mu + sigma * g.normal());
    uint64_t s = (uint64_t)std::llround(v > 1e18 ? 1e18 : v);
    if (s < lo) s = lo;
    if (s > hi) s = hi;
    sizes[i] = s;
  }
}

}  // extern "C"
