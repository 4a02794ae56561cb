// zero::Codec and the MI355X codec behind it.
//
// The interface keeps the reference's shape (src/archive.h:45-69): compress() turns a batch of whole blocks into one
// BlockOut per block, decompress() turns a list of BlockIn into their raw bytes back to back, in memory the codec owns
// and keeps valid through the next call (the container checks and writes batch k-1 while batch k decodes). Additions,
// each optional for a codec and for a caller:
//   - block_crcs(), range_crcs(): the CRC-32 of every block of the last batch, and of byte ranges of it (the files'
//     pieces), computed where the batch already is, so the container need not checksum blocks or files on the host
//     (src/archive.h:279-287,376,618);
//   - batch_memory(): host memory the codec reads and writes fastest (pinned, for DMA), for the caller to fill;
//   - BlockOut::view: a payload left in codec-owned memory (valid through the next compress() and until the one
//     after, as decompress() output is) instead of copied into BlockOut::data, when the caller asks for views.
//
// GpuCodec replaces src/gpu_codec.h: it calls libzero_hip.so (include/zero_hip.h), whose kernels are compiled ahead of
// time for gfx950, so prepare() has nothing to build.
#pragma once
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "zero_hip.h"

namespace zero {

// damaged data or a format error (the reference's zb::Error, src/codec.h:29-31)
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// the GPU failed (the reference's gpu::Error, src/gpu.h:17-27)
struct GpuError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct BlockOut {
  uint8_t method = 0;         // 0 stored, 1 ZB2
  std::vector<uint8_t> data;  // the payload when method 1 (stored blocks are written from the input)
  const uint8_t* view = nullptr;  // or the payload in codec-owned memory (view_len bytes)
  uint32_t view_len = 0;
  const uint8_t* bytes() const { return view ? view : data.data(); }
  size_t size() const { return view ? view_len : data.size(); }
};
struct BlockIn {
  uint8_t method = 0;
  uint32_t raw_len = 0;
  const uint8_t* data = nullptr;  // comp_len bytes
  uint32_t comp_len = 0;
};

struct Codec {
  virtual ~Codec() = default;
  virtual std::string name() const = 0;
  // in: n bytes = whole blocks of 2^block_log (the last may be shorter); out gets one entry per block
  virtual void compress(const uint8_t* in, size_t n, uint32_t block_log, std::vector<BlockOut>& out) = 0;
  // the blocks' raw bytes back to back, valid through the next call and until the one after
  virtual const uint8_t* decompress(const std::vector<BlockIn>& blocks) = 0;
  virtual size_t batch_bytes(uint32_t block_log) const = 0;
  virtual size_t batch_bytes_decode(uint32_t block_log) const { return batch_bytes(block_log); }
  virtual void prepare(uint32_t) {}
  // CRC-32 of each block of the last compress() input or decompress() output; nullptr when not computed
  virtual const uint32_t* block_crcs() const { return nullptr; }
  // CRC-32 of byte ranges [off[i], off[i] + len[i]) of the last compress() input (input) or decompress() output,
  // computed where that data already is; false when the codec cannot (the container then checksums files itself)
  virtual bool range_crcs(bool input, const std::vector<uint64_t>& off, const std::vector<uint64_t>& len,
                          std::vector<uint32_t>& crc) {
    (void)input; (void)off; (void)len; (void)crc;
    return false;
  }
  // `bytes` of host memory for the caller's batch number `slot` (0..3), which compress() and decompress() read
  // fastest; nullptr: any memory serves as well
  virtual uint8_t* batch_memory(size_t bytes, int slot) { (void)bytes; (void)slot; return nullptr; }
  // payloads as views into codec-owned memory instead of copies, where the codec can
  virtual void hand_out_views(bool on) { (void)on; }
};

class GpuCodec : public Codec {
 public:
  int level = 5;  // 1 fast .. 9 thorough (src/gpu_codec.h:66-70)

  explicit GpuCodec(int ordinal = 0) {
    if (zgpu_open(ordinal, &g_) != ZGPU_OK || !g_) throw GpuError("no usable GPU (HIP)");
    zgpu_props(g_, &props_);
    // ZERO_FAIL_GPU=n: the GPU's batch n + 1 fails as a lost device would (src/safe_codec.h:4,17,39), for tests
    if (const char* e = getenv("ZERO_FAIL_GPU")) fail_after_ = atoi(e);
    if (const char* e = getenv("ZERO_PROFILE")) profile_ = atoi(e) != 0;
    // ZERO_BATCH_BYTES: a smaller batch (at least 64 KiB), so that small inputs run as many batches (for tests)
    if (const char* e = getenv("ZERO_BATCH_BYTES")) batch_ = std::max<size_t>((size_t)1 << 16, (size_t)atoll(e));
  }
  ~GpuCodec() override {
    for (Pinned* p : {&mem_[0], &mem_[1], &mem_[2], &mem_[3], &pay_[0], &pay_[1], &out_[0], &out_[1], &blob_}) p->release(g_);
    zgpu_close(g_);
  }
  GpuCodec(const GpuCodec&) = delete;
  GpuCodec& operator=(const GpuCodec&) = delete;

  const zgpu_props_t& props() const { return props_; }
  std::string name() const override {
    std::string n = props_.name[0] ? props_.name : "AMD GPU";
    return "GPU " + n + " (" + props_.arch + ")";
  }
  // 1 GiB per batch: 4096 blocks of 2^18 in flight, in chunks over two streams (about 19 bytes of device scratch per
  // input byte of a chunk)
  size_t batch_bytes(uint32_t) const override { return batch_; }
  size_t batch_bytes_decode(uint32_t) const override { return batch_; }
  void prepare(uint32_t block_log) override { block_log_ = block_log; }
  const uint32_t* block_crcs() const override { return crcs_.empty() ? nullptr : crcs_.data(); }
  bool range_crcs(bool input, const std::vector<uint64_t>& off, const std::vector<uint64_t>& len,
                  std::vector<uint32_t>& crc) override {
    crc.assign(off.size(), 0);
    check(zgpu_crc32_ranges(g_, input ? 0 : 1, (uint32_t)off.size(), off.data(), len.data(), crc.data()));
    return true;
  }
  uint8_t* batch_memory(size_t bytes, int slot) override {
    if (slot < 0 || slot > 3) return nullptr;
    return mem_[slot].get(g_, bytes);
  }
  void hand_out_views(bool on) override { views_ = on; }

  void compress(const uint8_t* in, size_t n, uint32_t block_log, std::vector<BlockOut>& out) override {
    const size_t B = (size_t)1 << block_log;
    const size_t nb = (n + B - 1) / B;
    out.assign(nb, BlockOut{});
    crcs_.assign(nb, 0);
    if (!nb) return;
    batch_fails();
    const double t0 = now();
    pay_at_ ^= 1;  // (the last batch's payload views stay valid: it may still be being written)
    uint8_t* pay = pay_[pay_at_].get(g_, n);
    lens_.resize(nb);
    check(zgpu_compress(g_, in, n, block_log, level, pay, n, lens_.data(), crcs_.data()));
    size_t o = 0;
    for (size_t b = 0; b < nb; ++b) {
      if (lens_[b] == ZGPU_STORED) continue;
      out[b].method = 1;
      if (views_) {
        out[b].view = pay + o;
        out[b].view_len = lens_[b];
      } else {
        out[b].data.assign(pay + o, pay + o + lens_[b]);
      }
      o += lens_[b];
    }
    profile_line("compress", n, t0);
  }

  const uint8_t* decompress(const std::vector<BlockIn>& blocks) override {
    const size_t nb = blocks.size();
    crcs_.assign(nb, 0);
    if (!nb) return nullptr;
    batch_fails();
    const double t0 = now();
    size_t total = 0, raw_total = 0;
    bool contiguous = true;  // (one archive read gives one blob)
    for (size_t b = 0; b < nb; ++b) {
      total += blocks[b].comp_len;
      raw_total += blocks[b].raw_len;
      if (b && blocks[b].data != blocks[b - 1].data + blocks[b - 1].comp_len) contiguous = false;
    }
    const uint8_t* comp = blocks[0].data;
    if (!contiguous) {
      uint8_t* blob = blob_.get(g_, total);
      size_t o = 0;
      for (const BlockIn& k : blocks) { memcpy(blob + o, k.data, k.comp_len); o += k.comp_len; }
      comp = blob;
    }
    clen_.resize(nb);
    rlen_.resize(nb);
    err_.assign(nb, 0);
    meth_.resize(nb);
    for (size_t b = 0; b < nb; ++b) {
      clen_[b] = blocks[b].comp_len;
      rlen_[b] = blocks[b].raw_len;
      meth_[b] = blocks[b].method;
    }
    out_at_ ^= 1;  // (the previous batch may still be being checked and written)
    uint8_t* raw = out_[out_at_].get(g_, raw_total + 1);
    const int rc = zgpu_decompress(g_, comp, clen_.data(), meth_.data(), rlen_.data(), (uint32_t)nb, block_log_, raw,
                                   err_.data(), crcs_.data());
    if (rc == ZGPU_ERR_DAMAGED)
      for (size_t b = 0; b < nb; ++b)
        if (err_[b]) throw Error("block damaged (GPU decoder, code " + std::to_string(err_[b]) + ")");
    check(rc);
    profile_line("decompress", raw_total, t0);
    return raw;
  }

 private:
  // pinned host memory that grows on demand
  struct Pinned {
    uint8_t* p = nullptr;
    size_t n = 0;
    uint8_t* get(zgpu_t* g, size_t bytes) {
      if (bytes <= n && p) return p;
      release(g);
      void* q = nullptr;
      if (zgpu_alloc_host(g, bytes ? bytes : 1, &q) != ZGPU_OK) throw GpuError(zgpu_last_error(g));
      p = (uint8_t*)q;
      n = bytes;
      return p;
    }
    void release(zgpu_t* g) {
      if (p) zgpu_free_host(g, p);
      p = nullptr;
      n = 0;
    }
  };
  void batch_fails() {
    if (fail_after_ == 0) throw GpuError("the GPU was reset (device lost, simulated by ZERO_FAIL_GPU)");
    if (fail_after_ > 0) --fail_after_;
  }
  void check(int rc) {
    if (rc == ZGPU_OK) return;
    const std::string m = zgpu_last_error(g_);
    if (rc == ZGPU_ERR_DAMAGED || rc == ZGPU_ERR_ARG) throw Error(m);
    throw GpuError(m);
  }
  // ZERO_PROFILE=1: a line per batch, in the reference's form (src/gpu_codec.h:32-49): the kernel stages of the
  // batch's first chunk by events on its stream, then the whole call
  void profile_line(const char* what, size_t bytes, double t0) {
    if (!profile_) return;
    const char* names[32];
    float ms[32];
    const int k = zgpu_stage_times(g_, names, ms, 32);
    std::string line;
    char b[64];
    for (int i = 0; i < k; ++i) {
      snprintf(b, sizeof b, " %s %.1f", names[i], ms[i]);
      line += b;
    }
    fprintf(stderr, "[profile] %s %.1f MB (ms):%s (first chunk) call %.1f\n", what, bytes / 1e6, line.c_str(),
            (now() - t0) * 1e3);
  }

  zgpu_t* g_ = nullptr;
  zgpu_props_t props_{};
  uint32_t block_log_ = 18;
  size_t batch_ = (size_t)1 << 30;
  int fail_after_ = -1;
  bool profile_ = false, views_ = false;
  Pinned mem_[4], pay_[2], out_[2], blob_;
  int pay_at_ = 0, out_at_ = 0;
  std::vector<uint32_t> lens_, crcs_, clen_, rlen_, err_;
  std::vector<uint8_t> meth_;
};

}  // namespace zero
