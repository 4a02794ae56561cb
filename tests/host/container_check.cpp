// The container's host logic (zero-compression_amd/host/archive.h) driven by a stand-in codec on the CPU: blocks are
// stored, and range_crcs() answers from the host copy of the last batch, the way the GPU codec answers from the
// device copy. Batches hold 3 blocks, so files and their pieces cross batch boundaries. Test infrastructure.
//   container_check c <out.zero> <block_log> <inputs...>   |   container_check t <archive>   |   x <archive> <dir>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "archive.h"

struct HostCodec : zero::Codec {
  std::vector<uint8_t> last_in, last_out;
  std::string name() const override { return "host stand-in"; }
  void compress(const uint8_t* in, size_t n, uint32_t block_log, std::vector<zero::BlockOut>& out) override {
    last_in.assign(in, in + n);
    const size_t B = (size_t)1 << block_log;
    out.assign((n + B - 1) / B, zero::BlockOut{});  // (every block stored)
  }
  const uint8_t* decompress(const std::vector<zero::BlockIn>& blocks) override {
    at ^= 1;  // (the batch before stays valid: the container writes it while this one decodes)
    std::vector<uint8_t>& o = outs[at];
    o.clear();
    for (const zero::BlockIn& b : blocks) {
      if (b.method != 0) throw zero::Error("the stand-in codec reads stored blocks only");
      o.insert(o.end(), b.data, b.data + b.comp_len);
    }
    last_out.assign(o.begin(), o.end());
    o.push_back(0);
    return o.data();
  }
  std::vector<uint8_t> outs[2];
  int at = 0;
  size_t batch_bytes(uint32_t block_log) const override { return (size_t)3 << block_log; }
  bool range_crcs(bool input, const std::vector<uint64_t>& off, const std::vector<uint64_t>& len,
                  std::vector<uint32_t>& crc) override {
    if (host_sums) return false;  // (the container then checksums the files' pieces itself)
    const std::vector<uint8_t>& d = input ? last_in : last_out;
    crc.resize(off.size());
    for (size_t i = 0; i < off.size(); ++i) {
      if (off[i] + len[i] > d.size()) throw zero::Error("range outside the batch");
      crc[i] = zero::crc::of(d.data() + off[i], len[i]);
    }
    ++calls;
    return true;
  }
  size_t calls = 0;
  bool host_sums = getenv("CHECK_HOST_SUMS") != nullptr;
};

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string mode = argv[1];
  HostCodec codec;
  zero::Report rep;
  try {
    if (mode == "c") {
      std::vector<std::string> in(argv + 4, argv + argc);
      zero::compress_archive(in, argv[2], codec, (uint32_t)std::stoul(argv[3]), rep);
    } else if (mode == "t") {
      zero::test_archive(argv[2], codec, rep);
    } else if (mode == "x") {
      zero::extract_archive(argv[2], argv[3], codec, rep);
    } else {
      return 2;
    }
  } catch (const std::exception& e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  printf("OK range_crc_calls=%zu\n", codec.calls);
  return 0;
}
