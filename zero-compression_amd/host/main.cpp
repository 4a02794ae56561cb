// zero: the command line, on the MI355X codec through libzero_hip.so.
//
//   zero c <file|folder>... -o <archive.zero> [--block N] [--level 1..9]
//   zero x <archive.zero> [-o <folder>]
//   zero l <archive.zero>
//   zero t <archive.zero>
//   zero bench <file> [--block N] [--level L]
//   zero info
//
// The commands, their defaults and the lines they print are the reference's (src/main.cpp:59-304), so scripts that
// parse them keep working. The codec is the GPU's: --cpu asks for the reference's CPU codec, which this program does
// not contain, and is refused. Environment: ZERO_PROFILE=1 prints where the time goes, ZERO_FAIL_GPU=n fails the
// GPU's batch n + 1 (for tests: the job stops and an unfinished archive is removed), ZERO_BENCH_REPEAT=k decodes k
// times in bench.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "archive.h"

#define ZERO_VERSION "0.1.0-mi355x"

namespace zero {
namespace cli {

// sizes as the reference prints them: bytes below 1 KiB, one decimal in KiB and MiB, two in GiB
std::string human(uint64_t b) {
  struct Unit {
    uint64_t below;
    double scale;
    const char* fmt;
  };
  static const Unit units[] = {{1ull << 10, 1.0, "%.0f B"},
                               {1ull << 20, 1024.0, "%.1f KiB"},
                               {1ull << 30, 1048576.0, "%.1f MiB"},
                               {~0ull, 1073741824.0, "%.2f GiB"}};
  const Unit* u = units;
  while (b >= u->below) ++u;
  char s[64];
  snprintf(s, sizeof s, u->fmt, (double)b / u->scale);
  return s;
}
std::string rate(uint64_t bytes, double secs) {
  if (secs <= 0) return "-";
  char s[64];
  snprintf(s, sizeof s, "%.0f MB/s", (double)bytes / secs / 1e6);
  return s;
}

struct Options {
  std::string cmd, out;
  std::vector<std::string> inputs;
  uint32_t block_log = 18;
  int level = 5;
  bool cpu = false, quiet = false;
};

void usage() {
  printf("zero %s: lossless compression of files and folders on the GPU (AMD MI355X, HIP)\n\n"
         "  zero c <file|folder>... -o <archive.zero> [--block N] [--level 1..9]\n"
         "  zero x <archive.zero> [-o <folder>]\n"
         "  zero l <archive.zero>\n"
         "  zero t <archive.zero>\n"
         "  zero bench <file> [--block N] [--level 1..9]\n"
         "  zero info\n\n"
         "  --block N   block size as log2 bytes, 16 (64 KiB) .. 20 (1 MiB); default 18 (256 KiB)\n"
         "  --level L   how hard to look for matches: 1 fast .. 9 thorough (default 5)\n"
         "  -q          no progress\n",
         ZERO_VERSION);
}

bool parse(int argc, char** argv, Options& o) {
  if (argc < 2) return false;
  o.cmd = argv[1];
  for (int i = 2; i < argc; ++i) {
    const std::string a = argv[i];
    const bool has_value = i + 1 < argc;
    if (a == "-o" && has_value) o.out = argv[++i];
    else if (a == "--block" && has_value) o.block_log = (uint32_t)atoi(argv[++i]);
    else if (a == "--level" && has_value) o.level = atoi(argv[++i]);
    else if (a == "--cpu") o.cpu = true;
    else if (a == "--gpu") o.cpu = false;
    else if (a == "-q") o.quiet = true;
    else if (a.size() > 1 && a[0] == '-') {
      fprintf(stderr, "unknown option %s\n", a.c_str());
      return false;
    } else {
      o.inputs.push_back(a);
    }
  }
  return true;
}

// "\r  42.0 %  1.2 GiB / 3.0 GiB" on stderr, at most ten times a second, and a newline at the end
class Progress {
 public:
  explicit Progress(bool quiet) : quiet_(quiet) {}
  void operator()(uint64_t done, uint64_t total) {
    if (quiet_) return;
    const double t = now();
    if (done < total && t - last_ < 0.1) return;
    last_ = t;
    fprintf(stderr, "\r  %5.1f %%  %s / %s      ", total ? 100.0 * (double)done / (double)total : 100.0,
            human(done).c_str(), human(total).c_str());
    if (done >= total) fputc('\n', stderr);
    fflush(stderr);
  }

 private:
  bool quiet_;
  double last_ = 0;
};

// the job's report, and its codec: the GPU
struct Job {
  Report rep;
  Progress progress;
  std::unique_ptr<GpuCodec> codec;
  explicit Job(const Options& o) : progress(o.quiet) {
    if (o.cpu) throw Error("--cpu: this build has no CPU codec (the MI355X is the codec)");
    rep.log = [](const std::string& s) { fprintf(stderr, "%s\n", s.c_str()); };
    rep.progress = [this](uint64_t d, uint64_t t) { progress(d, t); };
    codec = std::make_unique<GpuCodec>(0);
    codec->level = o.level;
  }
};

// the last part of a path, without trailing slashes (what an archive or folder is named after)
std::string last_part(std::string p) {
  while (p.size() > 1 && p.back() == '/') p.pop_back();
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? p : p.substr(s + 1);
}

int run_compress(const Options& o) {
  if (o.inputs.empty()) { fprintf(stderr, "nothing to compress\n"); return 2; }
  std::string out = o.out;
  if (out.empty()) {
    std::string base = last_part(o.inputs[0]);
    if (base == "." || base == ".." || base.empty()) base = last_part(tree_detail::input_name(o.inputs[0]));
    out = base + ".zero";
  }
  Job job(o);
  fprintf(stderr, "codec: %s, blocks of %s\n", job.codec->name().c_str(), human(1ull << o.block_log).c_str());
  const CompressStats st = compress_archive(o.inputs, out, *job.codec, o.block_log, job.rep);
  printf("%s: %s -> %s (%.1f %%), %llu blocks (%llu stored), %.2f s (%s; codec %.2f s, %s)\n", out.c_str(),
         human(st.raw).c_str(), human(st.compressed).c_str(), st.raw ? 100.0 * (double)st.compressed / (double)st.raw : 0.0,
         (unsigned long long)st.blocks, (unsigned long long)st.stored_blocks, st.seconds, rate(st.raw, st.seconds).c_str(),
         st.codec_seconds, rate(st.raw, st.codec_seconds).c_str());
  return 0;
}

int run_extract(const Options& o) {
  if (o.inputs.size() != 1) { fprintf(stderr, "give one archive\n"); return 2; }
  std::string out = o.out;
  if (out.empty()) {  // the archive's name without its extension, next to where we are
    const std::string name = last_part(o.inputs[0]);
    const size_t dot = name.find_last_of('.');
    out = dot == std::string::npos || dot == 0 ? name : name.substr(0, dot);
    if (out.empty() || out == o.inputs[0]) out = ".";
  }
  Job job(o);
  fprintf(stderr, "codec: %s\n", job.codec->name().c_str());
  const ExtractStats st = extract_archive(o.inputs[0], out, *job.codec, job.rep);
  printf("%s: %zu files, %zu folders, %s in %.2f s (%s; codec %.2f s, %s), all checksums match\n", out.c_str(),
         st.files, st.folders, human(st.raw).c_str(), st.seconds, rate(st.raw, st.seconds).c_str(), st.codec_seconds,
         rate(st.raw, st.codec_seconds).c_str());
  return 0;
}

int run_test(const Options& o) {
  if (o.inputs.size() != 1) { fprintf(stderr, "give one archive\n"); return 2; }
  Job job(o);
  const ExtractStats st = test_archive(o.inputs[0], *job.codec, job.rep);
  printf("%s: OK, %zu files, %zu folders, %s in %llu blocks, %.2f s (%s)\n", o.inputs[0].c_str(), st.files, st.folders,
         human(st.raw).c_str(), (unsigned long long)st.blocks, st.seconds, rate(st.raw, st.seconds).c_str());
  return 0;
}

int run_list(const Options& o) {
  if (o.inputs.size() != 1) { fprintf(stderr, "give one archive\n"); return 2; }
  const ArchiveInfo a = list_archive(o.inputs[0]);
  const uint64_t comp = a.foot.table_off - fmt::HEADER_BYTES - a.hdr.blocks * fmt::RECORD_BYTES;
  printf("%s: %zu entries, %s in %llu blocks of %s, compressed to %s (%.1f %%)\n", o.inputs[0].c_str(),
         a.entries.size(), human(a.hdr.raw_bytes).c_str(), (unsigned long long)a.hdr.blocks,
         human(1ull << a.hdr.block_log).c_str(), human(comp).c_str(),
         a.hdr.raw_bytes ? 100.0 * (double)comp / (double)a.hdr.raw_bytes : 0.0);
  for (const fmt::Item& e : a.entries) {
    if (e.folder) printf("%14s  %s/\n", "", e.name.c_str());
    else printf("%14llu  %s\n", (unsigned long long)e.size, e.name.c_str());
  }
  return 0;
}

int run_info(const Options&) {
  try {
    GpuCodec g(0);
    const zgpu_props_t& p = g.props();
    printf("GPU: %s\n", g.name().c_str());
    printf("  %d CUs, %s memory, work-group up to %u, local memory %u KiB, %d MHz\n", p.cu_count,
           human(p.hbm_bytes).c_str(), p.max_group, p.lds_per_block / 1024, p.clock_khz / 1000);
    printf("  kernels: ready (compiled ahead of time for gfx950)\n");
  } catch (const std::exception& e) {
    printf("GPU: not available: %s\n", e.what());
  }
  printf("CPU: %u threads\n", std::thread::hardware_concurrency());
  return 0;
}

// one file compressed in memory in the codec's batches, then decompressed and compared; the speeds of both
int run_bench(const Options& o) {
  if (o.inputs.size() != 1) { fprintf(stderr, "give one file\n"); return 2; }
  std::vector<uint8_t> data;
  {
    const int fd = ::open(o.inputs[0].c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) { fprintf(stderr, "cannot open %s\n", o.inputs[0].c_str()); return 1; }
    struct stat sb;
    fstat(fd, &sb);
    data.resize((size_t)sb.st_size);
    size_t got = 0;
    while (got < data.size()) {
      const ssize_t r = ::read(fd, data.data() + got, data.size() - got);
      if (r <= 0) break;
      got += (size_t)r;
    }
    ::close(fd);
    if (got != data.size()) { fprintf(stderr, "read failed\n"); return 1; }
  }
  Job job(o);
  Codec& c = *job.codec;
  const uint64_t B = 1ull << o.block_log;
  printf("%s: %s, blocks of %s, codec %s\n", o.inputs[0].c_str(), human(data.size()).c_str(), human(B).c_str(),
         c.name().c_str());
  c.prepare(o.block_log);  // (before the clock starts)
  const size_t batch = (size_t)std::max<uint64_t>(B, c.batch_bytes(o.block_log) / B * B);
  // compress: each batch's payloads kept for the decode
  std::vector<BlockIn> ins;
  std::vector<std::vector<uint8_t>> payloads;  // (an inner vector keeps its bytes where they are when this one grows)
  std::vector<BlockOut> outs;
  uint64_t comp = 0, stored = 0;
  double tc = 0;
  for (size_t at = 0; at < data.size(); at += batch) {
    const size_t n = std::min(batch, data.size() - at);
    const double t0 = now();
    c.compress(data.data() + at, n, o.block_log, outs);
    tc += now() - t0;
    for (size_t b = 0; b < outs.size(); ++b) {
      BlockIn in;
      in.raw_len = (uint32_t)std::min<uint64_t>(B, n - b * B);
      if (outs[b].method == 0 || outs[b].size() >= in.raw_len) {
        in.method = 0;
        in.data = data.data() + at + b * B;
        in.comp_len = in.raw_len;
        ++stored;
      } else {
        payloads.emplace_back(outs[b].bytes(), outs[b].bytes() + outs[b].size());
        in.method = 1;
        in.data = payloads.back().data();
        in.comp_len = (uint32_t)outs[b].size();
      }
      comp += in.comp_len;
      ins.push_back(in);
    }
  }
  printf("  compress:   %.3f s, %s, %s -> %s (%.2f %%), %llu blocks, %llu stored\n", tc, rate(data.size(), tc).c_str(),
         human(data.size()).c_str(), human(comp).c_str(), data.empty() ? 0.0 : 100.0 * (double)comp / (double)data.size(),
         (unsigned long long)ins.size(), (unsigned long long)stored);
  const size_t per = (size_t)std::max<uint64_t>(1, c.batch_bytes_decode(o.block_log) / B);
  std::vector<uint8_t> back(data.size());
  bool all_ok = true;
  const int repeat = getenv("ZERO_BENCH_REPEAT") ? std::max(1, atoi(getenv("ZERO_BENCH_REPEAT"))) : 1;
  for (int r = 0; r < repeat; ++r) {
    const double t0 = now();
    size_t o2 = 0;
    for (size_t i = 0; i < ins.size(); i += per) {
      const std::vector<BlockIn> part(ins.begin() + i, ins.begin() + std::min(ins.size(), i + per));
      const uint8_t* raw = c.decompress(part);
      size_t n = 0;
      for (const BlockIn& k : part) n += k.raw_len;
      memcpy(back.data() + o2, raw, n);
      o2 += n;
    }
    const double td = now() - t0;
    const bool ok = back == data;
    all_ok = all_ok && ok;
    printf("  decompress: %.3f s, %s with %s: %s\n", td, rate(data.size(), td).c_str(), c.name().c_str(),
           ok ? "identical" : "MISMATCH");
  }
  return all_ok ? 0 : 1;
}

int main(int argc, char** argv) {
  Options o;
  if (!parse(argc, argv, o)) { usage(); return 2; }
  const double t_start = now();
  struct Teardown {
    double t0;
    ~Teardown() {
      if (getenv("ZERO_PROFILE")) fprintf(stderr, "[profile] main, up to teardown %.3f s\n", now() - t0);
    }
  } teardown{t_start};
  static const struct {
    const char* name;
    int (*run)(const Options&);
  } commands[] = {{"c", run_compress}, {"a", run_compress}, {"x", run_extract}, {"t", run_test},
                  {"l", run_list},     {"bench", run_bench}, {"info", run_info}};
  if (o.cmd == "-h" || o.cmd == "--help" || o.cmd == "help") { usage(); return 0; }
  for (const auto& c : commands) {
    if (o.cmd != c.name) continue;
    try {
      return c.run(o);
    } catch (const Cancelled&) {
      fprintf(stderr, "cancelled\n");
      return 3;
    } catch (const std::exception& e) {
      fprintf(stderr, "zero: %s\n", e.what());
      return 1;
    }
  }
  usage();
  return 2;
}

}  // namespace cli
}  // namespace zero

int main(int argc, char** argv) { return zero::cli::main(argc, argv); }
