// Compressing files and folders into a .zero archive, and listing, testing and extracting one, over any zero::Codec
// (src/archive.h:334-785 is the reference's container; the byte layout, in format.h, is the same).
//
// Both directions run as a pipeline over three batch slots: while the codec works on batch k, one worker thread
// reads batch k + 1 (from the files, or from the archive) and another checks and writes batch k - 1. A slot is
// reused for batch k + 3 only after batch k's write is done, and the codec's own buffers (payload views, decoded
// bytes) stay valid through the call after the one that filled them, which the write of batch k needs while batch
// k + 1 is in the codec. Checksums come from the codec where it can compute them (block_crcs, range_crcs: on the
// GPU, where the batch is), else from the host.
#pragma once
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <future>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "codec.h"
#include "crc32.h"
#include "format.h"
#include "solid.h"
#include "tree.h"

namespace zero {

// what the user interface hears while an operation runs
struct Report {
  std::function<void(const std::string&)> log;
  std::function<void(uint64_t done, uint64_t total)> progress;  // bytes of the solid stream done
  std::atomic<bool>* cancel = nullptr;
  void say(const std::string& s) const { if (log) log(s); }
  void at(uint64_t d, uint64_t t) const { if (progress) progress(d, t); }
  bool cancelled() const { return cancel && cancel->load(); }
};
struct Cancelled : std::runtime_error {
  Cancelled() : std::runtime_error("cancelled") {}
};

struct CompressStats {
  uint64_t raw = 0, compressed = 0, blocks = 0, stored_blocks = 0;
  size_t files = 0, folders = 0;
  double seconds = 0, codec_seconds = 0;
};
struct ExtractStats {
  uint64_t raw = 0, compressed = 0, blocks = 0;
  size_t files = 0, folders = 0;
  double seconds = 0, codec_seconds = 0;
  double read_seconds = 0, check_seconds = 0, write_seconds = 0;  // (worker threads' busy time)
};

// a thread that runs tasks in the order they are given
class Worker {
 public:
  Worker() : t_([this] { loop(); }) {}
  ~Worker() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_one();
    t_.join();
  }
  std::future<void> run(std::function<void()> f) {
    std::packaged_task<void()> task(std::move(f));
    std::future<void> fut = task.get_future();
    {
      std::lock_guard<std::mutex> lk(mu_);
      q_.push_back(std::move(task));
    }
    cv_.notify_one();
    return fut;
  }

 private:
  void loop() {
    for (;;) {
      std::packaged_task<void()> task;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
        if (q_.empty()) return;
        task = std::move(q_.front());
        q_.pop_front();
      }
      task();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::packaged_task<void()>> q_;
  bool stop_ = false;
  std::thread t_;
};

// waits for every future that is still pending (after a failure, so nothing runs on after the pipeline is gone)
inline void drain(std::vector<std::future<void>*> fs) {
  for (std::future<void>* f : fs)
    if (f->valid()) {
      try { f->get(); } catch (...) {}
    }
}

// host memory for a batch: the codec's (pinned) when it has some, else our own
struct SlotMemory {
  std::vector<uint8_t> own;
  uint8_t* p = nullptr;
  uint8_t* get(Codec& c, size_t bytes, int slot) {
    if (uint8_t* q = c.batch_memory(bytes, slot)) return p = q;
    if (own.size() < bytes) own.resize(bytes);
    return p = own.data();
  }
};

// the archive being written: appended through a large buffer; removed unless committed (an archive left
// unfinished by an error or a cancel could only fail later, src/archive.h:347-352)
class ArchiveFile {
 public:
  explicit ArchiveFile(const std::string& path) : path_(path) {
    fd_ = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (fd_ < 0) throw Error("cannot create " + path);
    buf_.reserve(CAP);
  }
  ~ArchiveFile() {
    if (fd_ >= 0) {
      ::close(fd_);
      ::unlink(path_.c_str());
    }
  }
  void put(const void* p, size_t n) {
    const uint8_t* s = (const uint8_t*)p;
    if (buf_.size() + n > CAP) flush();
    if (n >= CAP) { write_all(s, n); return; }
    buf_.insert(buf_.end(), s, s + n);
  }
  void commit() {
    flush();
    if (::close(fd_) != 0) throw Error("writing " + path_ + " failed (disk full?)");
    fd_ = -1;
  }

 private:
  static constexpr size_t CAP = (size_t)8 << 20;
  void flush() {
    write_all(buf_.data(), buf_.size());
    buf_.clear();
  }
  void write_all(const uint8_t* p, size_t n) {
    while (n) {
      const ssize_t w = ::write(fd_, p, n);
      if (w < 0 && errno == EINTR) continue;
      if (w <= 0) throw Error("writing " + path_ + " failed (disk full?)");
      p += w;
      n -= (size_t)w;
    }
  }
  std::string path_;
  int fd_ = -1;
  std::vector<uint8_t> buf_;
};

// the pieces of stream range [a, a + n) as offsets within that range (for Codec::range_crcs), and their entries
struct Pieces {
  std::vector<uint64_t> off, len;
  std::vector<size_t> entry;
  void of(const Layout& lay, uint64_t a, uint64_t n) {
    off.clear(), len.clear(), entry.clear();
    lay.pieces(a, n, [&](size_t i, uint64_t o, uint64_t l) {
      off.push_back(lay.start(i) + o - a);
      len.push_back(l);
      entry.push_back(i);
    });
  }
};

// ---------------------------------------------------------------- compress
inline CompressStats compress_archive(const std::vector<std::string>& inputs, const std::string& out_path, Codec& codec,
                                      uint32_t block_log, Report& rep) {
  if (block_log < fmt::MIN_BLOCK_LOG || block_log > fmt::MAX_BLOCK_LOG)
    throw Error("block size must be 64 KiB .. 1 MiB (log2 16..20)");
  const double t0 = now();
  CompressStats st;
  const std::vector<Source> src = scan_inputs(inputs, [&](const std::string& s) { rep.say(s); });
  std::vector<fmt::Item> items;
  for (const Source& s : src) {
    items.push_back(s.item);
    if (s.item.folder) ++st.folders; else ++st.files;
  }
  const Layout lay(items);
  const uint64_t total = lay.total();
  st.raw = total;
  rep.say(std::to_string(st.files) + " files, " + std::to_string(st.folders) + " folders, " + std::to_string(total) +
          " bytes");
  const uint64_t B = 1ull << block_log;
  ArchiveFile arc(out_path);
  fmt::Header hdr;
  hdr.block_log = block_log;
  hdr.raw_bytes = total;
  hdr.blocks = (total + B - 1) / B;
  uint8_t h[fmt::HEADER_BYTES];
  fmt::put(hdr, h);
  arc.put(h, sizeof h);

  const uint64_t batch = std::max<uint64_t>(B, codec.batch_bytes(block_log) / B * B);
  const uint64_t nbatch = (total + batch - 1) / batch;
  std::vector<uint32_t> probe;
  const bool codec_file_sums = codec.range_crcs(true, {}, {}, probe);
  codec.hand_out_views(true);
  FileSums sums(items.size());
  StreamReader reader(src, lay);
  struct Slot {
    SlotMemory mem;
    uint64_t at = 0;
    size_t n = 0;
    std::vector<BlockOut> outs;
    std::vector<uint32_t> crcs;
    Pieces pieces;
    std::vector<uint32_t> piece_crcs;  // (empty: computed on the host when written)
  } slot[3];
  Worker disk_in, disk_out;
  std::future<void> reading[3], writing[3];
  auto start_read = [&](uint64_t k) {
    Slot& s = slot[k % 3];
    s.at = k * batch;
    s.n = (size_t)std::min<uint64_t>(batch, total - s.at);
    uint8_t* m = s.mem.get(codec, (size_t)batch, (int)(k % 3));
    reading[k % 3] = disk_in.run([&reader, &s, m] { reader.read(s.at, m, s.n); });
  };
  auto write_batch = [&](Slot& s) {  // (the archive's blocks go out in batch order: one writer)
    const uint8_t* in = s.mem.p;
    for (size_t b = 0; b < s.outs.size(); ++b) {
      const size_t off = b * B;
      const uint32_t len = (uint32_t)std::min<uint64_t>(B, s.n - off);
      const BlockOut& o = s.outs[b];
      const bool stored = o.method == 0 || o.size() >= len;
      fmt::Record r;
      r.raw_len = len;
      r.comp_len = stored ? len : (uint32_t)o.size();
      r.method = stored ? fmt::METHOD_STORED : o.method;
      r.crc = s.crcs.empty() ? crc::of(in + off, len) : s.crcs[b];
      uint8_t rb[fmt::RECORD_BYTES];
      fmt::put(r, rb);
      arc.put(rb, sizeof rb);
      arc.put(stored ? in + off : o.bytes(), r.comp_len);
      st.compressed += r.comp_len;
      st.stored_blocks += stored;
      ++st.blocks;
    }
    for (size_t i = 0; i < s.pieces.off.size(); ++i)
      sums.add(s.pieces.entry[i], s.piece_crcs.empty() ? crc::of(in + s.pieces.off[i], s.pieces.len[i]) : s.piece_crcs[i],
               s.pieces.len[i]);
    rep.at(s.at + s.n, total);
  };
  rep.at(0, total);
  try {
    if (nbatch) start_read(0);
    for (uint64_t k = 0; k < nbatch; ++k) {
      if (rep.cancelled()) throw Cancelled();
      Slot& s = slot[k % 3];
      reading[k % 3].get();
      // batch k - 2 is written: its slot is free, and so are the codec buffers it used (this call reuses them)
      if (writing[(k + 1) % 3].valid()) writing[(k + 1) % 3].get();
      if (k + 1 < nbatch) start_read(k + 1);
      const double c0 = now();
      codec.compress(s.mem.p, s.n, block_log, s.outs);
      const uint32_t* bc = codec.block_crcs();
      s.crcs.assign(bc ? bc : nullptr, bc ? bc + s.outs.size() : nullptr);
      s.pieces.of(lay, s.at, s.n);
      s.piece_crcs.clear();
      if (codec_file_sums && !s.pieces.off.empty() && !codec.range_crcs(true, s.pieces.off, s.pieces.len, s.piece_crcs))
        throw Error("the codec lost its checksum support");
      st.codec_seconds += now() - c0;
      writing[k % 3] = disk_out.run([&write_batch, &s] { write_batch(s); });
    }
    for (auto& w : writing)
      if (w.valid()) w.get();
  } catch (...) {
    drain({&reading[0], &reading[1], &reading[2], &writing[0], &writing[1], &writing[2]});
    throw;
  }
  // the table, its place, and the footer
  for (size_t i = 0; i < items.size(); ++i)
    if (!items[i].folder) items[i].crc = items[i].size ? sums.of(i) : 0;
  const std::vector<uint8_t> tbl = fmt::table_bytes(items);
  arc.put(tbl.data(), tbl.size());
  fmt::Footer ft;
  ft.table_off = fmt::HEADER_BYTES + st.compressed + st.blocks * fmt::RECORD_BYTES;
  ft.table_len = tbl.size();
  ft.table_crc = crc::of(tbl.data(), tbl.size());
  uint8_t fb[fmt::FOOTER_BYTES];
  fmt::put(ft, fb);
  arc.put(fb, sizeof fb);
  arc.commit();
  st.seconds = now() - t0;
  return st;
}

// ---------------------------------------------------------------- reading
struct ArchiveInfo {
  fmt::Header hdr;
  fmt::Footer foot;
  uint64_t file_size = 0;
  std::vector<fmt::Item> entries;
};

// an archive opened for reading, with positional reads
class ArchiveIn {
 public:
  explicit ArchiveIn(const std::string& path) : path_(path) {
    fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd_ < 0) throw Error("cannot open " + path);
  }
  ~ArchiveIn() { ::close(fd_); }
  void read(uint64_t at, void* p, size_t n, const char* what) const {
    size_t got = 0;
    while (got < n) {
      const ssize_t r = ::pread(fd_, (uint8_t*)p + got, n - got, (off_t)(at + got));
      if (r < 0 && errno == EINTR) continue;
      if (r <= 0) fmt::damaged(what);
      got += (size_t)r;
    }
  }
  // header, footer and table, each checked
  ArchiveInfo info() const {
    ArchiveInfo a;
    struct stat sb;
    if (fstat(fd_, &sb) != 0) throw Error("cannot read " + path_);
    a.file_size = (uint64_t)sb.st_size;
    uint8_t h[fmt::HEADER_BYTES];
    if (a.file_size < fmt::HEADER_BYTES) throw Error("not a .zero archive");
    read(0, h, sizeof h, "header");
    a.hdr = fmt::get_header(h);
    if (a.file_size < fmt::HEADER_BYTES + fmt::FOOTER_BYTES) fmt::damaged("too short");
    uint8_t f[fmt::FOOTER_BYTES];
    read(a.file_size - fmt::FOOTER_BYTES, f, sizeof f, "footer (the file was cut short?)");
    a.foot = fmt::get_footer(f);
    if (a.foot.table_off < fmt::HEADER_BYTES || a.foot.table_len > a.file_size ||
        a.foot.table_off + a.foot.table_len + fmt::FOOTER_BYTES != a.file_size)
      fmt::damaged("table position");
    std::vector<uint8_t> t((size_t)a.foot.table_len);
    read(a.foot.table_off, t.data(), t.size(), "table cut");
    if (crc::of(t.data(), t.size()) != a.foot.table_crc) fmt::damaged("table checksum");
    a.entries = fmt::parse_table(t.data(), t.size(), a.hdr.raw_bytes);
    return a;
  }

 private:
  std::string path_;
  int fd_ = -1;
};

inline ArchiveInfo list_archive(const std::string& path) { return ArchiveIn(path).info(); }

// Decodes every block in order and hands the solid stream to sink(stream offset, bytes, n) on the writer thread,
// batch by batch; every block's CRC and every file's CRC is checked.
inline ExtractStats decode_archive(const std::string& path, Codec& codec, Report& rep, ArchiveInfo& a,
                                   const std::function<void(uint64_t, const uint8_t*, size_t)>& sink) {
  const double t0 = now();
  ExtractStats st;
  const ArchiveIn arc(path);
  a = arc.info();
  codec.prepare(a.hdr.block_log);
  const uint64_t B = 1ull << a.hdr.block_log;
  const uint64_t per = std::max<uint64_t>(1, codec.batch_bytes_decode(a.hdr.block_log) / B);
  const uint64_t nbatch = (a.hdr.blocks + per - 1) / per;
  const Layout lay(a.entries);
  FileSums sums(a.entries.size());
  std::vector<uint32_t> probe;
  const bool codec_file_sums = codec.range_crcs(false, {}, {}, probe);
  struct Slot {
    SlotMemory mem;
    std::vector<BlockIn> ins;
    std::vector<uint32_t> want;  // the records' CRCs
    uint64_t at = 0, raw = 0, comp = 0, first = 0;
    const uint8_t* out = nullptr;
    std::vector<uint32_t> crcs, piece_crcs;
    Pieces pieces;
  } slot[3];
  uint64_t next_pos = fmt::HEADER_BYTES, next_at = 0;  // (the reader's place in the archive and in the stream)
  Worker disk_in, disk_out;
  std::future<void> reading[3], writing[3];
  double t_read = 0, t_check = 0, t_write = 0;
  auto read_batch = [&](Slot& s, uint64_t first, uint64_t nb) {
    const double r0 = now();
    s.ins.assign(nb, BlockIn{});
    s.want.resize(nb);
    s.first = first;
    s.at = next_at;
    // the records first (16 bytes each, in front of their payloads), then the payloads in one read per block into
    // the slot's memory, which is at most the batch's raw size plus its records (a payload is smaller than its
    // block or stored)
    uint8_t* m = s.mem.get(codec, (size_t)(per * B), (int)(&s - slot));
    uint64_t pos = next_pos, used = 0, raw = 0;
    for (uint64_t b = 0; b < nb; ++b) {
      uint8_t rb[fmt::RECORD_BYTES];
      if (pos + fmt::RECORD_BYTES > a.foot.table_off) fmt::damaged("block header cut");
      arc.read(pos, rb, sizeof rb, "block header cut");
      pos += fmt::RECORD_BYTES;
      const fmt::Record r = fmt::get_record(rb, a.hdr.block_log);
      if (r.comp_len > a.foot.table_off - pos) fmt::damaged("block past the table");
      if (used + r.comp_len > per * B) fmt::damaged("block header");
      arc.read(pos, m + used, r.comp_len, "block cut");
      s.ins[b].method = r.method;
      s.ins[b].raw_len = r.raw_len;
      s.ins[b].comp_len = r.comp_len;
      s.ins[b].data = m + used;
      s.want[b] = r.crc;
      pos += r.comp_len;
      used += r.comp_len;
      raw += r.raw_len;
    }
    if (raw > a.hdr.raw_bytes - next_at) fmt::damaged("the blocks hold more bytes than the table lists");
    s.raw = raw;
    s.comp = used;
    next_pos = pos;
    next_at += raw;
    t_read += now() - r0;
  };
  auto check_and_write = [&](Slot& s) {
    const double c1 = now();
    uint64_t o = 0;
    for (size_t b = 0; b < s.ins.size(); ++b) {
      const uint32_t got = s.crcs.empty() ? crc::of(s.out + o, s.ins[b].raw_len) : s.crcs[b];
      if (got != s.want[b])
        throw Error("block " + std::to_string(s.first + b) + " does not match its checksum: the archive is damaged");
      o += s.ins[b].raw_len;
    }
    for (size_t i = 0; i < s.pieces.off.size(); ++i)
      sums.add(s.pieces.entry[i],
               s.piece_crcs.empty() ? crc::of(s.out + s.pieces.off[i], s.pieces.len[i]) : s.piece_crcs[i],
               s.pieces.len[i]);
    const double c2 = now();
    sink(s.at, s.out, (size_t)s.raw);
    t_check += c2 - c1;
    t_write += now() - c2;
    st.blocks += s.ins.size();
    st.raw += s.raw;
    rep.at(st.raw, a.hdr.raw_bytes);
  };
  rep.at(0, a.hdr.raw_bytes);
  auto start_read = [&](uint64_t k) {
    Slot& s = slot[k % 3];
    const uint64_t first = k * per, nb = std::min<uint64_t>(per, a.hdr.blocks - first);
    reading[k % 3] = disk_in.run([&read_batch, &s, first, nb] { read_batch(s, first, nb); });
  };
  try {
    if (nbatch) start_read(0);
    for (uint64_t k = 0; k < nbatch; ++k) {
      if (rep.cancelled()) throw Cancelled();
      Slot& s = slot[k % 3];
      reading[k % 3].get();
      if (writing[(k + 1) % 3].valid()) writing[(k + 1) % 3].get();  // (as in compress_archive)
      if (k + 1 < nbatch) start_read(k + 1);
      const double c0 = now();
      s.out = codec.decompress(s.ins);
      const uint32_t* bc = codec.block_crcs();
      s.crcs.assign(bc ? bc : nullptr, bc ? bc + s.ins.size() : nullptr);
      s.pieces.of(lay, s.at, s.raw);
      s.piece_crcs.clear();
      if (codec_file_sums && !s.pieces.off.empty() && !codec.range_crcs(false, s.pieces.off, s.pieces.len, s.piece_crcs))
        throw Error("the codec lost its checksum support");
      st.codec_seconds += now() - c0;
      st.compressed += s.comp;
      writing[k % 3] = disk_out.run([&check_and_write, &s] { check_and_write(s); });
    }
    for (auto& w : writing)
      if (w.valid()) w.get();
  } catch (...) {
    drain({&reading[0], &reading[1], &reading[2], &writing[0], &writing[1], &writing[2]});
    throw;
  }
  if (next_pos != a.foot.table_off) fmt::damaged("the blocks do not end where the table starts");
  if (st.raw != a.hdr.raw_bytes)
    fmt::damaged("the blocks hold " + std::to_string(st.raw) + " bytes, the table " + std::to_string(a.hdr.raw_bytes));
  for (size_t i = 0; i < a.entries.size(); ++i) {
    const fmt::Item& e = a.entries[i];
    if (!e.folder && e.size > 0 && sums.of(i) != e.crc) throw Error("file " + e.name + " does not match its checksum");
  }
  st.read_seconds = t_read;
  st.check_seconds = t_check;
  st.write_seconds = t_write;
  st.seconds = now() - t0;
  return st;
}

inline ExtractStats extract_archive(const std::string& path, const std::string& out_dir, Codec& codec, Report& rep) {
  ArchiveInfo a;
  std::unique_ptr<Layout> lay;
  std::unique_ptr<StreamWriter> w;
  ExtractStats st = decode_archive(path, codec, rep, a, [&](uint64_t at, const uint8_t* p, size_t n) {
    if (!w) {
      lay = std::make_unique<Layout>(a.entries);
      w = std::make_unique<StreamWriter>(a.entries, *lay, out_dir);
    }
    w->write(at, p, n);
  });
  if (!w) {
    lay = std::make_unique<Layout>(a.entries);
    w = std::make_unique<StreamWriter>(a.entries, *lay, out_dir);
  }
  w->finish();
  st.files = w->files();
  st.folders = w->folders();
  if (getenv("ZERO_PROFILE"))
    fprintf(stderr, "[profile] extract (s): read %.3f codec %.3f check %.3f write %.3f, whole %.3f\n", st.read_seconds,
            st.codec_seconds, st.check_seconds, st.write_seconds, st.seconds);
  return st;
}

// decodes everything and checks every block's and every file's checksum, writing nothing
inline ExtractStats test_archive(const std::string& path, Codec& codec, Report& rep) {
  ArchiveInfo a;
  ExtractStats st = decode_archive(path, codec, rep, a, [](uint64_t, const uint8_t*, size_t) {});
  for (const fmt::Item& e : a.entries) {
    if (e.folder) ++st.folders; else ++st.files;
  }
  return st;
}

}  // namespace zero
