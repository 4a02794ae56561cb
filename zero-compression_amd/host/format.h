// The .zero archive's byte layout, as plain values and their 24/16/24-byte encodings. Every integer is little endian.
//
//   header   "ZERO", u8 version 2, u8 block_log 16..20, u16 flags 0, u64 raw bytes, u64 blocks
//   block    u32 raw_len, u32 comp_len, u8 method (0 stored, 1 ZB2), 3 zero bytes, u32 CRC-32 of the raw bytes,
//            then comp_len bytes of payload
//   table    u32 entries; each: u8 kind (1 folder, else file), 3 zero bytes, u32 name length, the UTF-8 name
//            ('/' between parts, relative), u64 size, i64 mtime (Unix seconds), u32 attributes, u32 CRC-32
//   footer   u64 table offset, u64 table length, u32 CRC-32 of the table, "ZEND"
//
// The reference writes and reads exactly these bytes (src/archive.h:3-11,356-361,395-401,416-435,452-501,576-577;
// its comments say version 1 and "ZB1", its code writes version 2 and reads only that). The files' bytes, in table
// order, form one stream cut into blocks of 2^block_log bytes (the last one shorter); each block decodes alone.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "codec.h"
#include "crc32.h"

namespace zero {
namespace fmt {

constexpr size_t HEADER_BYTES = 24, RECORD_BYTES = 16, FOOTER_BYTES = 24;
constexpr uint8_t VERSION = 2;
constexpr uint32_t MIN_BLOCK_LOG = 16, MAX_BLOCK_LOG = 20;
constexpr uint8_t METHOD_STORED = 0, METHOD_ZB2 = 1;

[[noreturn]] inline void damaged(const std::string& what) { throw Error("archive damaged: " + what); }

inline void le(uint8_t* p, uint64_t v, int bytes) {
  for (int i = 0; i < bytes; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
inline uint64_t le(const uint8_t* p, int bytes) {
  uint64_t v = 0;
  for (int i = bytes - 1; i >= 0; --i) v = v << 8 | p[i];
  return v;
}

struct Header {
  uint32_t block_log = 18;
  uint64_t raw_bytes = 0, blocks = 0;
};
inline void put(const Header& h, uint8_t* o) {
  memcpy(o, "ZERO", 4);
  o[4] = VERSION;
  o[5] = (uint8_t)h.block_log;
  le(o + 6, 0, 2);
  le(o + 8, h.raw_bytes, 8);
  le(o + 16, h.blocks, 8);
}
inline Header get_header(const uint8_t* in) {
  if (memcmp(in, "ZERO", 4) != 0) throw Error("not a .zero archive");
  if (in[4] != VERSION)
    throw Error("archive version " + std::to_string(in[4]) + ": this program reads version " + std::to_string(VERSION));
  Header h;
  h.block_log = in[5];
  h.raw_bytes = le(in + 8, 8);
  h.blocks = le(in + 16, 8);
  if (h.block_log < MIN_BLOCK_LOG || h.block_log > MAX_BLOCK_LOG) damaged("block size");
  const uint64_t B = 1ull << h.block_log;
  if (h.blocks != (h.raw_bytes + B - 1) / B) damaged("block count");
  return h;
}

struct Record {
  uint32_t raw_len = 0, comp_len = 0;
  uint8_t method = METHOD_STORED;
  uint32_t crc = 0;
};
inline void put(const Record& r, uint8_t* o) {
  le(o, r.raw_len, 4);
  le(o + 4, r.comp_len, 4);
  o[8] = r.method;
  o[9] = o[10] = o[11] = 0;
  le(o + 12, r.crc, 4);
}
inline Record get_record(const uint8_t* in, uint32_t block_log) {
  Record r;
  r.raw_len = (uint32_t)le(in, 4);
  r.comp_len = (uint32_t)le(in + 4, 4);
  r.method = in[8];
  r.crc = (uint32_t)le(in + 12, 4);
  if (r.raw_len > (1u << block_log) || r.method > METHOD_ZB2 ||
      (r.method == METHOD_STORED && r.comp_len != r.raw_len))
    damaged("block header");
  return r;
}

struct Footer {
  uint64_t table_off = 0, table_len = 0;
  uint32_t table_crc = 0;
};
inline void put(const Footer& f, uint8_t* o) {
  le(o, f.table_off, 8);
  le(o + 8, f.table_len, 8);
  le(o + 16, f.table_crc, 4);
  memcpy(o + 20, "ZEND", 4);
}
inline Footer get_footer(const uint8_t* in) {
  if (memcmp(in + 20, "ZEND", 4) != 0) damaged("footer (the file was cut short?)");
  Footer f;
  f.table_off = le(in, 8);
  f.table_len = le(in + 8, 8);
  f.table_crc = (uint32_t)le(in + 16, 4);
  return f;
}

// one entry of the table
struct Item {
  bool folder = false;
  std::string name;
  uint64_t size = 0;
  int64_t mtime = 0;
  uint32_t attr = 0, crc = 0;
};

inline std::vector<uint8_t> table_bytes(const std::vector<Item>& items) {
  size_t n = 4;
  for (const Item& it : items) n += 32 + it.name.size();
  std::vector<uint8_t> t(n, 0);
  uint8_t* p = t.data();
  le(p, items.size(), 4);
  p += 4;
  for (const Item& it : items) {
    p[0] = it.folder ? 1 : 0;
    le(p + 4, it.name.size(), 4);
    memcpy(p + 8, it.name.data(), it.name.size());
    p += 8 + it.name.size();
    le(p, it.size, 8);
    le(p + 8, (uint64_t)it.mtime, 8);
    le(p + 16, it.attr, 4);
    le(p + 20, it.crc, 4);
    p += 24;
  }
  return t;
}

// the table's entries; their file sizes must add up to raw_bytes
inline std::vector<Item> parse_table(const uint8_t* t, size_t n, uint64_t raw_bytes) {
  if (n < 4) damaged("table");
  const uint8_t* p = t + 4;
  const uint8_t* end = t + n;
  const uint64_t count = le(t, 4);
  std::vector<Item> items;
  items.reserve((size_t)std::min<uint64_t>(count, n / 32));
  uint64_t files = 0;
  for (uint64_t i = 0; i < count; ++i) {
    if (end - p < 8) damaged("entry");
    Item it;
    it.folder = p[0] == 1;
    const uint64_t nl = le(p + 4, 4);
    p += 8;
    if ((uint64_t)(end - p) < nl + 24) damaged("entry name");
    it.name.assign((const char*)p, (size_t)nl);
    p += nl;
    it.size = le(p, 8);
    it.mtime = (int64_t)le(p + 8, 8);
    it.attr = (uint32_t)le(p + 16, 4);
    it.crc = (uint32_t)le(p + 20, 4);
    p += 24;
    if (it.folder) it.size = 0;
    if (it.size > raw_bytes - files) damaged("sizes");
    files += it.size;
    items.push_back(std::move(it));
  }
  if (files != raw_bytes) damaged("sizes");
  return items;
}

// A name from an archive may only name something inside the folder it is extracted to: relative, '/' between
// non-empty parts, none of them "." or "..", and no '\' or ':' (a drive or another separator elsewhere).
inline bool name_is_safe(const std::string& name) {
  if (name.empty() || name.front() == '/' || name.back() == '/') return false;
  if (name.find_first_of("\\:") != std::string::npos) return false;
  size_t a = 0;
  while (a <= name.size()) {
    size_t b = name.find('/', a);
    if (b == std::string::npos) b = name.size();
    const size_t len = b - a;
    if (len == 0 || (len == 1 && name[a] == '.') || (len == 2 && name[a] == '.' && name[a + 1] == '.')) return false;
    a = b + 1;
  }
  return true;
}

}  // namespace fmt
}  // namespace zero
