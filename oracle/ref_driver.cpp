// A thin extern "C" driver over the reference's own CPU codec headers (src/codec.h, src/codec2.h, src/crc32.h),
// compiled from where they lie under /root/reference by oracle/Makefile into oracle/_ref/libzeroref.so.
// TEST INFRASTRUCTURE ONLY: the golden-vector script, the parity tests and bench.py's cpu_baseline leg load it.
// No reference source is copied here: this file only calls the reference functions.
//
// ref_cpu_compress / ref_cpu_decompress reproduce CpuCodec (src/archive.h:97-136): one zb2 block per task,
// spread over `threads` threads, the block stored when its payload is not smaller (src/archive.h:394).
#include "codec2.h"
#include "crc32.h"

#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

namespace {
template <class F>
void run_parallel(size_t n, int threads, const F& f) {
  std::atomic<size_t> next{0};
  std::atomic<bool> failed{false};
  auto body = [&] {
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= n || failed.load()) return;
      try { f(i); } catch (...) { failed = true; return; }
    }
  };
  const int nt = threads < 1 ? 1 : (int)std::min<size_t>((size_t)threads, n ? n : 1);
  std::vector<std::thread> ts;
  for (int t = 1; t < nt; ++t) ts.emplace_back(body);
  body();
  for (auto& t : ts) t.join();
  if (failed) throw zb::Error("failed");
}
}  // namespace

extern "C" {

uint32_t ref_crc32(const uint8_t* p, size_t n) { return crc32::of(p, n); }
uint32_t ref_crc32_combine(uint32_t a, uint32_t b, uint64_t len_b) { return crc32::combine(a, b, len_b); }
void ref_normalize(const uint32_t* counts, uint32_t total, uint16_t* freq) { zb::normalize(counts, total, freq); }

size_t ref_encode_stream(const uint8_t* sym, uint32_t n, uint8_t* out) {
  std::vector<uint8_t> v;
  zb::encode_stream(sym, n, v);
  memcpy(out, v.data(), v.size());
  return v.size();
}

long ref_decode_stream(const uint8_t* p, size_t avail, uint32_t n, uint8_t* sym) {
  try { return (long)zb::decode_stream(p, p + avail, n, sym); } catch (const std::exception&) { return -1; }
}

// payload size; *stored = 1 when zb2::compress_block says the block is better stored
size_t ref_compress_block(const uint8_t* in, uint32_t n, int depth, uint8_t* out, size_t cap, int* stored) {
  zb2::Params P;
  P.depth = depth;
  std::vector<uint8_t> v;
  *stored = zb2::compress_block(in, n, v, P) ? 0 : 1;
  if (v.size() > cap) return (size_t)-1;
  memcpy(out, v.data(), v.size());
  return v.size();
}

int ref_decompress_block(const uint8_t* in, size_t in_len, uint8_t* out, uint32_t n) {
  try { zb2::decompress_block(in, in_len, out, n); return 0; } catch (const std::exception&) { return -1; }
}

// CpuCodec::compress over n bytes: comp_len/method per block, the block bodies (payload, or the raw bytes when
// stored) back to back in out. Returns the bytes written, or -1 when out is too small.
long ref_cpu_compress(const uint8_t* in, size_t n, uint32_t block_log, int depth, int threads, uint8_t* out,
                      size_t out_cap, uint32_t* comp_len, uint8_t* method) {
  const size_t B = (size_t)1 << block_log;
  const size_t nb = (n + B - 1) / B;
  std::vector<std::vector<uint8_t>> pay(nb);
  run_parallel(nb, threads, [&](size_t b) {
    const size_t off = b * B;
    const uint32_t len = (uint32_t)std::min(B, n - off);
    zb2::Params P;
    P.depth = depth;
    if (zb2::compress_block(in + off, len, pay[b], P)) method[b] = 1;
    else method[b] = 0;
  });
  size_t o = 0;
  for (size_t b = 0; b < nb; ++b) {
    const size_t off = b * B;
    const uint32_t len = (uint32_t)std::min(B, n - off);
    const bool st = method[b] == 0 || pay[b].size() >= len;
    const uint8_t* src = st ? in + off : pay[b].data();
    const size_t cl = st ? len : pay[b].size();
    if (o + cl > out_cap) return -1;
    memcpy(out + o, src, cl);
    comp_len[b] = (uint32_t)cl;
    method[b] = st ? 0 : 1;
    o += cl;
  }
  return (long)o;
}

// CpuCodec::decompress: blocks back to back in comp, their raw bytes back to back into out. 0, or -1 on damage.
int ref_cpu_decompress(const uint8_t* comp, const uint32_t* comp_len, const uint8_t* method, const uint32_t* raw_len,
                       size_t nb, int threads, uint8_t* out) {
  std::vector<size_t> co(nb + 1, 0), ro(nb + 1, 0);
  for (size_t b = 0; b < nb; ++b) { co[b + 1] = co[b] + comp_len[b]; ro[b + 1] = ro[b] + raw_len[b]; }
  try {
    run_parallel(nb, threads, [&](size_t b) {
      if (method[b] == 0) {
        if (comp_len[b] != raw_len[b]) throw zb::Error("stored block with a wrong length");
        memcpy(out + ro[b], comp + co[b], raw_len[b]);
      } else {
        zb2::decompress_block(comp + co[b], comp_len[b], out + ro[b], raw_len[b]);
      }
    });
  } catch (const std::exception&) {
    return -1;
  }
  return 0;
}

}  // extern "C"
