// The .zero container (format: src/archive.h:1-11; pipelines: src/archive.h:170-785), over any zero::Codec.
//
//   header   24 B: "ZERO", u8 version 2, u8 block_log (16..20), u16 0, u64 total raw bytes, u64 blocks
//   block    u32 raw_len, u32 comp_len, u8 method (0 stored, 1 ZB2), 3 x u8 0, u32 CRC-32 of the raw bytes, payload
//   table    u32 entries; each: u8 type (1 folder), 3 x u8 0, u32 name length, UTF-8 name ('/'), u64 size,
//            i64 mtime (Unix seconds), u32 attributes (0), u32 CRC-32 of the file
//   footer   24 B: u64 table offset, u64 table length, u32 CRC-32 of the table, "ZEND"
//
// The files' bytes, in entry order, form one solid stream cut into blocks; every block decompresses on its own.
// Block checksums come from the codec when it computes them (the GPU does), else from the host.
#pragma once
#include <sys/stat.h>
#include <fcntl.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <functional>
#include <future>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "codec.h"
#include "crc32.h"

namespace zero {

namespace fs = std::filesystem;
constexpr uint32_t MIN_BLOCK_LOG = 16, MAX_BLOCK_LOG = 20;

inline double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Report {
  std::function<void(const std::string&)> log;
  std::function<void(uint64_t done, uint64_t total)> progress;
  std::atomic<bool>* cancel = nullptr;
  void say(const std::string& s) const { if (log) log(s); }
  void at(uint64_t d, uint64_t t) const { if (progress) progress(d, t); }
  bool cancelled() const { return cancel && cancel->load(); }
};
struct Cancelled : std::runtime_error {
  Cancelled() : std::runtime_error("cancelled") {}
};

struct Entry {
  bool dir = false;
  std::string name;  // relative, '/' separated
  uint64_t size = 0;
  int64_t mtime = 0;
  uint32_t attr = 0;
  uint32_t crc = 0;
  fs::path source;
};

inline int64_t mtime_of(const fs::path& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 ? (int64_t)st.st_mtim.tv_sec : 0;
}
inline void set_mtime(const fs::path& p, int64_t secs) {
  if (!secs) return;
  struct timespec ts[2];
  ts[0].tv_sec = ts[1].tv_sec = (time_t)secs;
  ts[0].tv_nsec = ts[1].tv_nsec = 0;
  ::utimensat(AT_FDCWD, p.c_str(), ts, 0);
}

// a file is one entry under its own name; a folder is itself and everything in it, sorted inside each folder;
// symbolic links are not followed; a name already taken keeps its first entry (src/archive.h:168-240)
inline std::vector<Entry> collect(const std::vector<std::string>& inputs, Report& rep) {
  std::vector<Entry> es;
  std::vector<std::string> seen;
  auto add = [&](Entry e) {
    if (std::find(seen.begin(), seen.end(), e.name) != seen.end()) {
      rep.say("skipped (same name already in the archive): " + e.name);
      return;
    }
    seen.push_back(e.name);
    es.push_back(std::move(e));
  };
  std::function<void(const fs::path&, const std::string&)> walk = [&](const fs::path& dir, const std::string& rel) {
    std::vector<fs::directory_entry> items;
    std::error_code ec;
    for (fs::directory_iterator it(dir, fs::directory_options::skip_permission_denied, ec), end; !ec && it != end;
         it.increment(ec))
      items.push_back(*it);
    std::sort(items.begin(), items.end(), [](const fs::directory_entry& a, const fs::directory_entry& b) {
      return a.path().filename().string() < b.path().filename().string();
    });
    for (const auto& it : items) {
      if (rep.cancelled()) throw Cancelled();
      const std::string name = rel + "/" + it.path().filename().string();
      std::error_code e2;
      if (it.is_symlink(e2)) continue;
      if (it.is_directory(e2)) {
        Entry d;
        d.dir = true;
        d.name = name;
        d.mtime = mtime_of(it.path());
        d.source = it.path();
        add(std::move(d));
        walk(it.path(), name);
      } else if (it.is_regular_file(e2)) {
        Entry f;
        f.name = name;
        f.size = it.file_size(e2);
        if (e2) f.size = 0;
        f.mtime = mtime_of(it.path());
        f.source = it.path();
        add(std::move(f));
      }
    }
  };
  for (const std::string& in : inputs) {
    std::error_code ec;
    fs::path p = fs::absolute(fs::path(in), ec);
    while (!p.empty() && !p.has_filename()) p = p.parent_path();
    if (!fs::exists(p, ec)) throw Error("not found: " + in);
    const std::string top = p.filename().string();
    Entry e;
    e.name = top;
    e.mtime = mtime_of(p);
    e.source = p;
    if (fs::is_directory(p, ec)) {
      e.dir = true;
      add(std::move(e));
      walk(p, top);
    } else if (fs::is_regular_file(p, ec)) {
      e.size = fs::file_size(p, ec);
      add(std::move(e));
    } else {
      rep.say("skipped (not a file or folder): " + in);
    }
  }
  return es;
}

inline void put32(std::vector<uint8_t>& o, uint32_t v) { for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i))); }
inline void put64(std::vector<uint8_t>& o, uint64_t v) { put32(o, (uint32_t)v); put32(o, (uint32_t)(v >> 32)); }
inline uint32_t get32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
inline uint64_t get64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }

struct CompressStats {
  uint64_t raw = 0, compressed = 0, blocks = 0, stored_blocks = 0;
  size_t files = 0, folders = 0;
  double seconds = 0, codec_seconds = 0;
};

// Each file's CRC-32 from the CRCs of its pieces in the batches of the solid stream, which the codec computes where
// the batch already is (on the GPU): crc(A||B) = combine(crc(A), crc(B), |B|). The reference checksums files on the
// host as it reads and writes them (src/archive.h:279-287,320,707,774). finished(i, crc) is called for every file
// once all its bytes have gone by, in entry order (empty files with CRC 0).
struct FileCrcs {
  const std::vector<Entry>& es;
  size_t idx = 0;         // the entry the next byte of the stream belongs to
  uint64_t in_file = 0;   // its bytes already seen
  uint32_t crc = 0;       // their CRC
  explicit FileCrcs(const std::vector<Entry>& e) : es(e) {}
  template <class F> void settle(F&& finished) {
    while (idx < es.size()) {
      const Entry& e = es[idx];
      if (e.dir) { ++idx; continue; }
      if (in_file < e.size) return;
      finished(idx, e.size ? crc : 0u);
      ++idx;
      in_file = 0;
      crc = 0;
    }
  }
  // the next n bytes of the stream, the last compress() input (input) or decompress() output of the codec
  template <class F> void batch(Codec& codec, bool input, uint64_t n, F&& finished) {
    std::vector<uint64_t> off, len;
    uint64_t at = 0;
    for (size_t i = idx, d = in_file; at < n;) {  // (the pieces, as the fold below will meet them)
      if (i >= es.size()) throw Error("archive damaged: more data than the table lists");
      const Entry& e = es[i];
      if (e.dir || d >= e.size) { ++i; d = 0; continue; }
      const uint64_t take = std::min<uint64_t>(n - at, e.size - d);
      off.push_back(at);
      len.push_back(take);
      at += take;
      d += take;
    }
    std::vector<uint32_t> pc;
    if (!codec.range_crcs(input, off, len, pc)) throw Error("the codec lost its checksum support");
    settle(finished);
    for (size_t k = 0; k < off.size(); ++k) {
      crc = crc::combine(crc, pc[k], len[k]);
      in_file += len[k];
      settle(finished);
    }
  }
  template <class F> void close(F&& finished) { settle(finished); }
};

// the entries' bytes in order, in batches, with each file's CRC (src/archive.h:289-332) unless the codec computes
// those (host_crc false)
struct SolidReader {
  std::vector<Entry>& es;
  size_t idx = 0;
  uint64_t pos = 0;
  FILE* f = nullptr;
  uint32_t crc = 0;
  bool host_crc = true;
  explicit SolidReader(std::vector<Entry>& e) : es(e) {}
  ~SolidReader() { if (f) fclose(f); }
  size_t read(uint8_t* buf, size_t n) {
    size_t got = 0;
    while (got < n && idx < es.size()) {
      Entry& e = es[idx];
      if (e.dir || pos >= e.size) {
        if (!e.dir && host_crc) e.crc = crc;
        if (f) { fclose(f); f = nullptr; }
        crc = 0;
        ++idx;
        pos = 0;
        continue;
      }
      if (!f && !(f = fopen(e.source.c_str(), "rb"))) throw Error("cannot read " + e.source.string());
      const size_t want = (size_t)std::min<uint64_t>(n - got, e.size - pos);
      const size_t r = fread(buf + got, 1, want, f);
      if (r != want) memset(buf + got + r, 0, want - r);  // (a file that shrank: zeros keep the table right)
      if (host_crc) crc = crc::update(crc, buf + got, want);
      got += want;
      pos += want;
    }
    return got;
  }
  void finish() {
    if (host_crc && idx < es.size() && !es[idx].dir && pos >= es[idx].size) es[idx].crc = crc;
    if (f) { fclose(f); f = nullptr; }
    idx = es.size();
  }
};

inline CompressStats compress_archive(const std::vector<std::string>& inputs, const std::string& out_path, Codec& codec,
                                      uint32_t block_log, Report& rep) {
  if (block_log < MIN_BLOCK_LOG || block_log > MAX_BLOCK_LOG) throw Error("block size must be 64 KiB .. 1 MiB (log2 16..20)");
  const double t0 = now();
  CompressStats st;
  std::vector<Entry> es = collect(inputs, rep);
  uint64_t total = 0;
  for (const Entry& e : es) { total += e.size; if (e.dir) ++st.folders; else ++st.files; }
  st.raw = total;
  rep.say(std::to_string(st.files) + " files, " + std::to_string(st.folders) + " folders, " + std::to_string(total) + " bytes");
  FILE* f = fopen(out_path.c_str(), "wb");
  if (!f) throw Error("cannot create " + out_path);
  struct Closer {  // (an archive left unfinished is removed)
    FILE* f;
    std::string p;
    ~Closer() { if (f) { fclose(f); std::remove(p.c_str()); } }
  } closer{f, out_path};
  const size_t B = (size_t)1 << block_log;
  std::vector<uint8_t> hdr;
  hdr.insert(hdr.end(), {'Z', 'E', 'R', 'O', 2, (uint8_t)block_log, 0, 0});
  put64(hdr, total);
  put64(hdr, (total + B - 1) / B);
  fwrite(hdr.data(), 1, hdr.size(), f);
  size_t batch = std::max(B, codec.batch_bytes(block_log) / B * B);
  std::vector<uint8_t> buf(batch), next_buf(batch);
  SolidReader reader(es);
  FileCrcs files(es);
  std::vector<uint32_t> probe;
  const bool codec_crc = codec.range_crcs(true, {}, {}, probe);  // (the files' CRCs on the GPU when it can)
  reader.host_crc = !codec_crc;
  auto set_crc = [&es](size_t i, uint32_t c) { es[i].crc = c; };
  uint64_t done = 0, queued = 0;
  std::vector<BlockOut> outs;
  rep.at(0, total);
  auto fill = [&reader, batch](std::vector<uint8_t>& b, uint64_t left) {
    return reader.read(b.data(), (size_t)std::min<uint64_t>(batch, left));
  };
  size_t n = fill(buf, total);
  queued = n;
  std::vector<uint32_t> host_crc;
  while (n) {
    if (rep.cancelled()) throw Cancelled();
    std::future<size_t> ahead;  // the next batch is read while the GPU works on this one
    if (queued < total) ahead = std::async(std::launch::async, fill, std::ref(next_buf), total - queued);
    const double c0 = now();
    codec.compress(buf.data(), n, block_log, outs);
    if (codec_crc) files.batch(codec, true, n, set_crc);
    st.codec_seconds += now() - c0;
    const uint32_t* crcs = codec.block_crcs();
    if (!crcs) {
      host_crc.resize(outs.size());
      for (size_t b = 0; b < outs.size(); ++b) host_crc[b] = crc::of(buf.data() + b * B, std::min(B, n - b * B));
      crcs = host_crc.data();
    }
    for (size_t b = 0; b < outs.size(); ++b) {
      const size_t off = b * B;
      const uint32_t len = (uint32_t)std::min(B, n - off);
      const BlockOut& o = outs[b];
      const bool stored = o.method == 0 || o.data.size() >= len;
      std::vector<uint8_t> bh;
      put32(bh, len);
      put32(bh, stored ? len : (uint32_t)o.data.size());
      bh.insert(bh.end(), {(uint8_t)(stored ? 0 : o.method), 0, 0, 0});
      put32(bh, crcs[b]);
      fwrite(bh.data(), 1, bh.size(), f);
      if (stored) { fwrite(buf.data() + off, 1, len, f); ++st.stored_blocks; st.compressed += len; }
      else { fwrite(o.data.data(), 1, o.data.size(), f); st.compressed += o.data.size(); }
      ++st.blocks;
    }
    done += n;
    rep.at(done, total);
    n = 0;
    if (ahead.valid()) {
      n = ahead.get();
      queued += n;
      buf.swap(next_buf);
    }
  }
  reader.finish();
  if (codec_crc) files.close(set_crc);
  std::vector<uint8_t> tbl;
  put32(tbl, (uint32_t)es.size());
  for (const Entry& e : es) {
    tbl.insert(tbl.end(), {(uint8_t)(e.dir ? 1 : 0), 0, 0, 0});
    put32(tbl, (uint32_t)e.name.size());
    tbl.insert(tbl.end(), e.name.begin(), e.name.end());
    put64(tbl, e.size);
    put64(tbl, (uint64_t)e.mtime);
    put32(tbl, e.attr);
    put32(tbl, e.crc);
  }
  std::vector<uint8_t> foot;
  put64(foot, 24 + st.compressed + st.blocks * 16);
  put64(foot, tbl.size());
  put32(foot, crc::of(tbl.data(), tbl.size()));
  foot.insert(foot.end(), {'Z', 'E', 'N', 'D'});
  fwrite(tbl.data(), 1, tbl.size(), f);
  fwrite(foot.data(), 1, foot.size(), f);
  if (fflush(f) != 0 || ferror(f)) throw Error("writing " + out_path + " failed (disk full?)");
  fclose(f);
  closer.f = nullptr;
  st.seconds = now() - t0;
  return st;
}

struct ArchiveInfo {
  uint32_t block_log = 0;
  uint64_t total_raw = 0, n_blocks = 0, table_off = 0, file_size = 0;
  std::vector<Entry> entries;
  bool codec_file_crcs = false;  // read_blocks checks the files' CRCs with the codec: sinks need not
};

// header, footer and entry table, every field checked (src/archive.h:452-501)
inline ArchiveInfo read_info(FILE* f) {
  ArchiveInfo a;
  uint8_t h[24];
  if (fread(h, 1, 24, f) != 24 || memcmp(h, "ZERO", 4) != 0) throw Error("not a .zero archive");
  if (h[4] != 2) throw Error("archive version " + std::to_string(h[4]) + ": this program reads version 2");
  a.block_log = h[5];
  a.total_raw = get64(h + 8);
  a.n_blocks = get64(h + 16);
  if (a.block_log < MIN_BLOCK_LOG || a.block_log > MAX_BLOCK_LOG) throw Error("archive damaged: block size");
  if (fseeko(f, 0, SEEK_END) != 0) throw Error("cannot seek");
  const int64_t size = (int64_t)ftello(f);
  if (size < 48) throw Error("archive damaged: too short");
  a.file_size = (uint64_t)size;
  uint8_t ft[24];
  fseeko(f, size - 24, SEEK_SET);
  if (fread(ft, 1, 24, f) != 24 || memcmp(ft + 20, "ZEND", 4) != 0) throw Error("archive damaged: footer (the file was cut short?)");
  a.table_off = get64(ft);
  const uint64_t tl = get64(ft + 8);
  const uint32_t tc = get32(ft + 16);
  if (a.table_off + tl + 24 != a.file_size) throw Error("archive damaged: table position");
  std::vector<uint8_t> tbl(tl);
  fseeko(f, (off_t)a.table_off, SEEK_SET);
  if (fread(tbl.data(), 1, tl, f) != tl) throw Error("archive damaged: table cut");
  if (crc::of(tbl.data(), tl) != tc) throw Error("archive damaged: table checksum");
  if (tl < 4) throw Error("archive damaged: table");
  const uint8_t* p = tbl.data();
  const uint8_t* end = p + tl;
  const uint32_t n = get32(p);
  p += 4;
  uint64_t sum = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (end - p < 8) throw Error("archive damaged: entry");
    Entry e;
    e.dir = p[0] == 1;
    const uint32_t nl = get32(p + 4);
    p += 8;
    if ((size_t)(end - p) < (size_t)nl + 24) throw Error("archive damaged: entry name");
    e.name.assign((const char*)p, nl);
    p += nl;
    e.size = get64(p);
    e.mtime = (int64_t)get64(p + 8);
    e.attr = get32(p + 16);
    e.crc = get32(p + 20);
    p += 24;
    if (!e.dir) sum += e.size;
    a.entries.push_back(std::move(e));
  }
  if (sum != a.total_raw) throw Error("archive damaged: sizes");
  return a;
}

inline ArchiveInfo list_archive(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw Error("cannot open " + path);
  struct Closer { FILE* f; ~Closer() { fclose(f); } } c{f};
  return read_info(f);
}

// a name from an archive is written only under the target folder (src/archive.h:510-522)
inline bool safe_name(const std::string& n) {
  if (n.empty() || n[0] == '/' || n.find('\\') != std::string::npos || n.find(':') != std::string::npos) return false;
  size_t i = 0;
  while (i <= n.size()) {
    size_t j = n.find('/', i);
    if (j == std::string::npos) j = n.size();
    const std::string part = n.substr(i, j - i);
    if (part.empty() || part == "." || part == "..") return false;
    i = j + 1;
  }
  return true;
}

struct ExtractStats {
  uint64_t raw = 0, compressed = 0, blocks = 0;
  size_t files = 0, folders = 0;
  double seconds = 0, codec_seconds = 0;
};

// every block decompressed in order, its CRC checked, and the solid stream handed to sink (src/archive.h:536-639):
// the next batch is read ahead while this one decodes, and the one before is checked and sunk meanwhile
inline ExtractStats read_blocks(const std::string& path, Codec& codec, Report& rep, ArchiveInfo& a,
                                const std::function<void(const uint8_t*, size_t)>& sink) {
  const double t0 = now();
  ExtractStats st;
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw Error("cannot open " + path);
  struct Closer { FILE* f; ~Closer() { fclose(f); } } c{f};
  a = read_info(f);
  codec.prepare(a.block_log);
  std::vector<uint32_t> probe;
  a.codec_file_crcs = codec.range_crcs(false, {}, {}, probe);  // (the files' CRCs on the GPU when it can)
  FileCrcs files(a.entries);
  auto check_crc = [&a](size_t i, uint32_t c) {
    if (a.entries[i].size > 0 && c != a.entries[i].crc)
      throw Error("file " + a.entries[i].name + " does not match its checksum");
  };
  fseeko(f, 24, SEEK_SET);
  const size_t B = (size_t)1 << a.block_log;
  const size_t per_batch = std::max(B, codec.batch_bytes_decode(a.block_log) / B * B) / B;
  struct Batch {
    std::vector<uint8_t> comp;
    std::vector<BlockIn> ins;
    std::vector<uint32_t> crcs;
    size_t raw_total = 0;
  };
  uint64_t done = 0, unread = a.n_blocks, undecoded = a.n_blocks, pos = 24;
  auto read_batch = [&](Batch& t, size_t nb) {
    t.comp.clear();
    t.ins.clear();
    t.crcs.clear();
    t.raw_total = 0;
    std::vector<size_t> offs;
    for (size_t b = 0; b < nb; ++b) {
      uint8_t h[16];
      if (fread(h, 1, 16, f) != 16) throw Error("archive damaged: block header cut");
      pos += 16;
      BlockIn in;
      in.raw_len = get32(h);
      in.comp_len = get32(h + 4);
      in.method = h[8];
      if (in.raw_len > B || (in.method == 0 && in.comp_len != in.raw_len) || in.method > 1) throw Error("archive damaged: block header");
      if (pos + in.comp_len > a.table_off) throw Error("archive damaged: block past the table");
      offs.push_back(t.comp.size());
      t.comp.resize(t.comp.size() + in.comp_len);
      if (fread(t.comp.data() + offs.back(), 1, in.comp_len, f) != in.comp_len) throw Error("archive damaged: block cut");
      pos += in.comp_len;
      t.crcs.push_back(get32(h + 12));
      t.raw_total += in.raw_len;
      t.ins.push_back(in);
    }
    for (size_t b = 0; b < nb; ++b) t.ins[b].data = t.comp.data() + offs[b];
  };
  rep.at(0, a.total_raw);
  Batch batches[2];
  int at = 0;
  std::future<void> reading, writing;
  auto take = [&](Batch& t) {
    const size_t nb = (size_t)std::min<uint64_t>(per_batch, unread);
    unread -= nb;
    return std::async(std::launch::async, read_batch, std::ref(t), nb);
  };
  if (unread) reading = take(batches[at]);
  while (undecoded > 0) {
    if (rep.cancelled()) throw Cancelled();
    reading.get();
    if (unread) reading = take(batches[at ^ 1]);
    Batch& cur = batches[at];
    const size_t nb = cur.ins.size();
    uint64_t comp_bytes = 0;
    for (const BlockIn& k : cur.ins) comp_bytes += k.comp_len;
    const double c0 = now();
    const uint8_t* raw = codec.decompress(cur.ins);
    st.codec_seconds += now() - c0;
    st.compressed += comp_bytes;
    if (writing.valid()) writing.get();  // (the batch before: its bytes stay valid until this call's next)
    std::vector<uint32_t> got(nb);
    if (const uint32_t* gc = codec.block_crcs()) std::copy(gc, gc + nb, got.begin());
    else {
      size_t o = 0;
      for (size_t b = 0; b < nb; ++b) { got[b] = crc::of(raw + o, cur.ins[b].raw_len); o += cur.ins[b].raw_len; }
    }
    for (size_t b = 0; b < nb; ++b)
      if (got[b] != cur.crcs[b])
        throw Error("block " + std::to_string(st.blocks + b) + " does not match its checksum: the archive is damaged");
    if (a.codec_file_crcs) files.batch(codec, false, cur.raw_total, check_crc);
    const size_t raw_total = cur.raw_total;
    writing = std::async(std::launch::async, [&, raw, raw_total, nb] {
      sink(raw, raw_total);
      done += raw_total;
      st.blocks += nb;
      rep.at(done, a.total_raw);
    });
    undecoded -= nb;
    at ^= 1;
  }
  if (writing.valid()) writing.get();
  if (done != a.total_raw) throw Error("archive damaged: the blocks hold " + std::to_string(done) + " bytes, the table " + std::to_string(a.total_raw));
  if (a.codec_file_crcs) files.close(check_crc);
  st.raw = done;
  st.seconds = now() - t0;
  return st;
}

// the solid stream written back into files under a folder (src/archive.h:642-722)
struct SolidWriter {
  const std::vector<Entry>& es;
  fs::path root;
  size_t idx = 0;
  uint64_t pos = 0;
  FILE* f = nullptr;
  uint32_t crc = 0;
  bool host_crc = true;  // (false: read_blocks checks the files' CRCs)
  size_t files = 0, folders = 0;
  SolidWriter(const std::vector<Entry>& e, const fs::path& r) : es(e), root(r) {}
  ~SolidWriter() { if (f) fclose(f); }
  void open_next() {
    const Entry& e = es[idx];
    if (!safe_name(e.name)) throw Error("refused: unsafe name in the archive: " + e.name);
    const fs::path p = root / fs::path(e.name);
    std::error_code ec;
    fs::create_directories(p.parent_path(), ec);
    if (!(f = fopen(p.c_str(), "wb"))) throw Error("cannot create " + p.string());
    ++files;
  }
  void finish_file() {
    const Entry& e = es[idx];
    if (f) { fclose(f); f = nullptr; }
    if (host_crc && !e.dir && crc != e.crc && e.size > 0) throw Error("file " + e.name + " does not match its checksum");
    set_mtime(root / fs::path(e.name), e.mtime);
    crc = 0;
  }
  void settle() {
    while (idx < es.size()) {
      const Entry& e = es[idx];
      if (e.dir) {
        if (!safe_name(e.name)) throw Error("refused: unsafe name in the archive: " + e.name);
        std::error_code ec;
        fs::create_directories(root / fs::path(e.name), ec);
        ++folders;
        ++idx;
        continue;
      }
      if (e.size == 0) { open_next(); finish_file(); ++idx; continue; }
      if (pos >= e.size) { finish_file(); ++idx; pos = 0; continue; }
      return;
    }
  }
  void write(const uint8_t* p, size_t n) {
    settle();
    while (n > 0) {
      if (idx >= es.size()) throw Error("archive damaged: more data than the table lists");
      const Entry& e = es[idx];
      if (!f) open_next();
      const size_t want = (size_t)std::min<uint64_t>(n, e.size - pos);
      if (fwrite(p, 1, want, f) != want) throw Error("writing " + e.name + " failed (disk full?)");
      if (host_crc) crc = crc::update(crc, p, want);
      p += want;
      n -= want;
      pos += want;
      settle();
    }
  }
  void close() {
    settle();
    for (const Entry& e : es)
      if (e.dir) set_mtime(root / fs::path(e.name), e.mtime);
  }
};

inline ExtractStats extract_archive(const std::string& path, const std::string& out_dir, Codec& codec, Report& rep) {
  ArchiveInfo a;
  const fs::path root(out_dir);
  std::error_code ec;
  fs::create_directories(root, ec);
  std::unique_ptr<SolidWriter> w;
  ExtractStats st = read_blocks(path, codec, rep, a, [&](const uint8_t* p, size_t n) {
    if (!w) { w = std::make_unique<SolidWriter>(a.entries, root); w->host_crc = !a.codec_file_crcs; }
    w->write(p, n);
  });
  if (!w) w = std::make_unique<SolidWriter>(a.entries, root);
  w->close();
  st.files = w->files;
  st.folders = w->folders;
  return st;
}

// everything decompressed and every block and file checksum checked, nothing written (src/archive.h:747-785)
inline ExtractStats test_archive(const std::string& path, Codec& codec, Report& rep) {
  ArchiveInfo a;
  size_t idx = 0, files = 0;
  uint64_t pos = 0;
  uint32_t crc = 0;
  auto settle = [&] {
    while (idx < a.entries.size()) {
      const Entry& e = a.entries[idx];
      if (e.dir) { ++idx; continue; }
      if (pos >= e.size) {
        if (!a.codec_file_crcs && e.size > 0 && crc != e.crc) throw Error("file " + e.name + " does not match its checksum");
        ++files;
        ++idx;
        pos = 0;
        crc = 0;
        continue;
      }
      return;
    }
  };
  ExtractStats st = read_blocks(path, codec, rep, a, [&](const uint8_t* p, size_t n) {
    settle();
    while (n > 0) {
      if (idx >= a.entries.size()) throw Error("archive damaged: more data than the table lists");
      const Entry& e = a.entries[idx];
      const size_t want = (size_t)std::min<uint64_t>(n, e.size - pos);
      if (!a.codec_file_crcs) crc = crc::update(crc, p, want);
      p += want;
      n -= want;
      pos += want;
      settle();
    }
  });
  settle();
  st.files = files;
  for (const Entry& e : a.entries)
    if (e.dir) ++st.folders;
  return st;
}

}  // namespace zero
