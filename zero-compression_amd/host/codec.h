// zero::Codec and the MI355X codec behind it.
//
// The interface keeps the reference's shape (src/archive.h:45-69): compress() turns a batch of whole blocks into one
// BlockOut per block, decompress() turns a list of BlockIn into their raw bytes back to back, in memory the codec owns
// and keeps valid through the next call (the container checks and writes batch k-1 while batch k decodes). The one
// additions are block_crcs() and range_crcs(): the CRC-32 of every block of the last batch, and of any byte ranges
// of it (the files' pieces), computed on the GPU where the batch already is, so the container need not checksum
// blocks or files on the host (src/archive.h:279-287,376,618).
//
// GpuCodec replaces src/gpu_codec.h: it calls libzero_hip.so (include/zero_hip.h), whose kernels are compiled ahead of
// time for gfx950, so prepare() has nothing to build.
#pragma once
#include <cstdint>
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

struct BlockOut {
  uint8_t method = 0;         // 0 stored, 1 ZB2
  std::vector<uint8_t> data;  // the payload when method 1 (stored blocks are written from the input)
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
};

class GpuCodec : public Codec {
 public:
  int level = 5;  // 1 fast .. 9 thorough (src/gpu_codec.h:66-70)

  explicit GpuCodec(int ordinal = 0) {
    if (zgpu_open(ordinal, &g_) != ZGPU_OK || !g_) throw GpuError("no usable GPU (HIP)");
    zgpu_props_t p;
    zgpu_props(g_, &p);
    props_ = p;
  }
  ~GpuCodec() override { zgpu_close(g_); }
  GpuCodec(const GpuCodec&) = delete;
  GpuCodec& operator=(const GpuCodec&) = delete;

  const zgpu_props_t& props() const { return props_; }
  std::string name() const override {
    std::string n = props_.name[0] ? props_.name : "AMD GPU";
    return "GPU " + n + " (" + props_.arch + ")";
  }
  // 1 GiB per batch: 4096 blocks of 2^18 in flight, about 19 bytes of device scratch per input byte
  size_t batch_bytes(uint32_t) const override { return (size_t)1 << 30; }
  size_t batch_bytes_decode(uint32_t) const override { return (size_t)1 << 30; }
  void prepare(uint32_t block_log) override { block_log_ = block_log; }
  const uint32_t* block_crcs() const override { return crcs_.empty() ? nullptr : crcs_.data(); }
  // on the GPU, from the batch the last call left there (zgpu_crc32_ranges)
  bool range_crcs(bool input, const std::vector<uint64_t>& off, const std::vector<uint64_t>& len,
                  std::vector<uint32_t>& crc) override {
    crc.assign(off.size(), 0);
    check(zgpu_crc32_ranges(g_, input ? 0 : 1, (uint32_t)off.size(), off.data(), len.data(), crc.data()));
    return true;
  }

  void compress(const uint8_t* in, size_t n, uint32_t block_log, std::vector<BlockOut>& out) override {
    const size_t B = (size_t)1 << block_log;
    const size_t nb = (n + B - 1) / B;
    out.assign(nb, BlockOut{});
    crcs_.assign(nb, 0);
    if (!nb) return;
    slabs_.resize(nb * B);
    lens_.resize(nb);
    check(zgpu_compress(g_, in, n, block_log, level, slabs_.data(), lens_.data(), crcs_.data()));
    for (size_t b = 0; b < nb; ++b) {
      if (lens_[b] == ZGPU_STORED) continue;
      out[b].method = 1;
      out[b].data.assign(slabs_.data() + b * B, slabs_.data() + b * B + lens_[b]);
    }
  }

  const uint8_t* decompress(const std::vector<BlockIn>& blocks) override {
    const size_t nb = blocks.size();
    if (!nb) return nullptr;
    size_t total = 0, raw_total = 0;
    for (const BlockIn& k : blocks) { total += k.comp_len; raw_total += k.raw_len; }
    bool contiguous = true;  // (one archive read gives one blob)
    for (size_t b = 1; b < nb && contiguous; ++b) contiguous = blocks[b].data == blocks[b - 1].data + blocks[b - 1].comp_len;
    const uint8_t* blob = blocks[0].data;
    if (!contiguous) {
      blob_.resize(total);
      size_t o = 0;
      for (const BlockIn& k : blocks) { memcpy(blob_.data() + o, k.data, k.comp_len); o += k.comp_len; }
      blob = blob_.data();
    }
    std::vector<uint32_t> clen(nb), rlen(nb), err(nb);
    std::vector<uint8_t> meth(nb);
    for (size_t b = 0; b < nb; ++b) { clen[b] = blocks[b].comp_len; rlen[b] = blocks[b].raw_len; meth[b] = blocks[b].method; }
    at_ ^= 1;  // (the previous batch may still be being checked and written)
    outbuf_[at_].resize(raw_total + 1);
    crcs_.assign(nb, 0);
    const int rc = zgpu_decompress(g_, blob, clen.data(), meth.data(), rlen.data(), (uint32_t)nb, block_log_,
                                   outbuf_[at_].data(), err.data(), crcs_.data());
    if (rc == ZGPU_ERR_DAMAGED) {
      for (size_t b = 0; b < nb; ++b)
        if (err[b]) throw Error("block damaged (GPU decoder, code " + std::to_string(err[b]) + ")");
    }
    check(rc);
    return outbuf_[at_].data();
  }

 private:
  void check(int rc) {
    if (rc == ZGPU_OK) return;
    const std::string m = zgpu_last_error(g_);
    if (rc == ZGPU_ERR_DAMAGED || rc == ZGPU_ERR_ARG) throw Error(m);
    throw GpuError(m);
  }
  zgpu_t* g_ = nullptr;
  zgpu_props_t props_{};
  uint32_t block_log_ = 18;
  std::vector<uint8_t> slabs_, blob_, outbuf_[2];
  std::vector<uint32_t> lens_, crcs_;
  int at_ = 0;
};

}  // namespace zero
