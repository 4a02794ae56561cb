// The solid stream: the files' bytes in table order, one after the other (folders hold none). A Layout maps stream
// offsets to (entry, offset in the entry) by binary search over the entries' start offsets, so any range of the
// stream, a batch of blocks in particular, splits into per-file pieces without a reader's or writer's state. Files
// are read and written with positional I/O (pread / pwrite) through a one-file descriptor cache, so a range can be
// handled wherever it falls.
//
// Behaviour the format asks for, as the reference has it (src/archive.h:289-332, 537-602, 643-785): a file that is
// shorter when read than when listed reads as zeros past its end (the stream keeps the length the table states);
// extraction makes every folder the table lists and every file's parent, sets modification times (folders last,
// since writing into a folder changes its time), and refuses a name that would leave the target folder.
#pragma once
#include <fcntl.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <string>
#include <vector>

#include "format.h"
#include "tree.h"

namespace zero {

class Layout {
 public:
  explicit Layout(const std::vector<fmt::Item>& items) : items_(items), start_(items.size() + 1, 0) {
    for (size_t i = 0; i < items.size(); ++i) start_[i + 1] = start_[i] + (items[i].folder ? 0 : items[i].size);
  }
  uint64_t total() const { return start_.back(); }
  uint64_t start(size_t i) const { return start_[i]; }
  size_t count() const { return items_.size(); }
  // f(entry, offset in the entry, length) for every file piece of the stream range [a, a + n), in order
  template <class F>
  void pieces(uint64_t a, uint64_t n, F&& f) const {
    const uint64_t e = a + n;
    // the first entry whose bytes reach past a
    size_t i = (size_t)(std::upper_bound(start_.begin(), start_.end(), a) - start_.begin());
    i = i ? i - 1 : 0;
    for (; i < items_.size() && start_[i] < e; ++i) {
      if (items_[i].folder || start_[i + 1] <= a) continue;
      const uint64_t lo = std::max(a, start_[i]), hi = std::min(e, start_[i + 1]);
      if (hi > lo) f(i, lo - start_[i], hi - lo);
    }
  }

 private:
  const std::vector<fmt::Item>& items_;
  std::vector<uint64_t> start_;  // stream offset of entry i's first byte; start_[count] = the total
};

// one open file at a time, kept open while consecutive pieces use it
class FileCache {
 public:
  enum Mode { READ, CREATE, APPEND };
  ~FileCache() { drop(); }
  int get(size_t entry, const std::string& path, Mode m) {
    if (fd_ >= 0 && entry_ == entry) return fd_;
    drop();
    const int flags = m == READ ? O_RDONLY : m == CREATE ? O_WRONLY | O_CREAT | O_TRUNC : O_WRONLY;
    fd_ = ::open(path.c_str(), flags | O_CLOEXEC, 0644);
    entry_ = entry;
    return fd_;
  }
  void drop() {
    if (fd_ >= 0) ::close(fd_);
    fd_ = -1;
    entry_ = (size_t)-1;
  }

 private:
  int fd_ = -1;
  size_t entry_ = (size_t)-1;
};

// the stream as the source files hold it
class StreamReader {
 public:
  StreamReader(const std::vector<Source>& src, const Layout& lay) : src_(src), lay_(lay) {}
  // fills buf with the stream range [a, a + n)
  void read(uint64_t a, uint8_t* buf, size_t n) {
    lay_.pieces(a, n, [&](size_t i, uint64_t off, uint64_t len) {
      uint8_t* dst = buf + (lay_.start(i) + off - a);
      const int fd = files_.get(i, src_[i].path, FileCache::READ);
      if (fd < 0) throw Error("cannot read " + src_[i].path);
      uint64_t got = 0;
      while (got < len) {
        const ssize_t r = ::pread(fd, dst + got, (size_t)(len - got), (off_t)(off + got));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) break;  // (the file shrank since it was listed)
        got += (uint64_t)r;
      }
      if (got < len) memset(dst + got, 0, (size_t)(len - got));
    });
  }

 private:
  const std::vector<Source>& src_;
  const Layout& lay_;
  FileCache files_;
};

inline void set_mtime(const std::string& path, int64_t secs) {
  if (!secs) return;
  struct timespec ts[2];
  ts[0].tv_sec = ts[1].tv_sec = (time_t)secs;
  ts[0].tv_nsec = ts[1].tv_nsec = 0;
  (void)utimensat(AT_FDCWD, path.c_str(), ts, 0);
}
inline void make_dirs(const std::string& path) {  // mkdir -p
  for (size_t k = 1; k <= path.size(); ++k)
    if (k == path.size() || path[k] == '/') (void)::mkdir(path.substr(0, k).c_str(), 0755);
}

// the stream written back into files under a root folder
class StreamWriter {
 public:
  StreamWriter(const std::vector<fmt::Item>& items, const Layout& lay, std::string root)
      : items_(items), lay_(lay), root_(std::move(root)), made_(items.size(), 0) {
    for (const fmt::Item& it : items)
      if (!fmt::name_is_safe(it.name)) throw Error("refused: unsafe name in the archive: " + it.name);
    make_dirs(root_);
    for (size_t i = 0; i < items.size(); ++i)
      if (items[i].folder) { make_dirs(path(i)); ++folders_; }
  }
  // the stream range [a, a + n) from p
  void write(uint64_t a, const uint8_t* p, size_t n) {
    lay_.pieces(a, n, [&](size_t i, uint64_t off, uint64_t len) {
      const int fd = open_file(i);
      const uint8_t* src = p + (lay_.start(i) + off - a);
      uint64_t put = 0;
      while (put < len) {
        const ssize_t w = ::pwrite(fd, src + put, (size_t)(len - put), (off_t)(off + put));
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) throw Error("writing " + items_[i].name + " failed (disk full?)");
        put += (uint64_t)w;
      }
    });
  }
  // the empty files, then the modification times (files, then folders)
  void finish() {
    for (size_t i = 0; i < items_.size(); ++i)
      if (!items_[i].folder && !made_[i]) open_file(i);
    files_.drop();
    for (size_t i = 0; i < items_.size(); ++i)
      if (!items_[i].folder) set_mtime(path(i), items_[i].mtime);
    for (size_t i = 0; i < items_.size(); ++i)
      if (items_[i].folder) set_mtime(path(i), items_[i].mtime);
  }
  size_t files() const { return files_made_; }
  size_t folders() const { return folders_; }

 private:
  std::string path(size_t i) const { return root_ + "/" + items_[i].name; }
  int open_file(size_t i) {
    const std::string p = path(i);
    if (!made_[i]) {
      const std::string parent = p.substr(0, p.find_last_of('/'));
      if (parent != last_parent_) {  // (a folder is made once, not again for every file in it)
        make_dirs(parent);
        last_parent_ = parent;
      }
    }
    const int fd = files_.get(i, p, made_[i] ? FileCache::APPEND : FileCache::CREATE);
    if (fd < 0) throw Error("cannot create " + p);
    if (!made_[i]) {
      made_[i] = 1;
      ++files_made_;
    }
    return fd;
  }
  const std::vector<fmt::Item>& items_;
  const Layout& lay_;
  std::string root_, last_parent_;
  std::vector<uint8_t> made_;
  FileCache files_;
  size_t files_made_ = 0, folders_ = 0;
};

// Every file's CRC-32 from the CRCs of its pieces, which arrive in stream order (a file's pieces in its order):
// crc(A || B) = combine(crc(A), crc(B), |B|).
class FileSums {
 public:
  explicit FileSums(size_t n) : crc_(n, 0) {}
  void add(size_t entry, uint32_t piece_crc, uint64_t len) { crc_[entry] = crc::combine(crc_[entry], piece_crc, len); }
  uint32_t of(size_t entry) const { return crc_[entry]; }

 private:
  std::vector<uint32_t> crc_;
};

}  // namespace zero
