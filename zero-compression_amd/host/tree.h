// What an archive holds for a list of inputs (the entries src/archive.h:189-240 produces, in the same order, so that
// the same inputs give the same archive): each input in turn; a file as one entry under its own name; a folder as
// itself and then everything under it, depth first, the entries of each folder in the byte order of their names.
// Symbolic links are not followed and not stored. A name already taken keeps its first entry (with a warning); a
// folder whose name was taken is still walked, so its contents join the first folder's.
//
// The walk is an explicit stack over POSIX directory reads: a folder's children are read and sorted once, and pushed
// in reverse so that they come off the stack in order, each folder just before its contents.
#pragma once
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdlib>
#include <functional>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

#include "format.h"

namespace zero {

struct Source {
  fmt::Item item;
  std::string path;  // on disk
};

namespace tree_detail {
inline bool stat_of(const std::string& path, struct stat& st) { return lstat(path.c_str(), &st) == 0; }
// an input's entry name: its last path part ("dir/" and "dir/." name dir; "." and ".." name the folder they are)
inline std::string input_name(std::string path) {
  while (path.size() > 1 && path.back() == '/') path.pop_back();
  std::string last = path.substr(path.find_last_of('/') == std::string::npos ? 0 : path.find_last_of('/') + 1);
  if (last == "." || last == ".." || last.empty()) {
    char* r = realpath(path.empty() ? "." : path.c_str(), nullptr);
    if (r) {
      std::string full = r;
      free(r);
      last = full.substr(full.find_last_of('/') + 1);
    }
  }
  return last;
}
}  // namespace tree_detail

inline std::vector<Source> scan_inputs(const std::vector<std::string>& inputs,
                                       const std::function<void(const std::string&)>& warn) {
  std::vector<Source> out;
  std::unordered_set<std::string> taken;
  struct Todo {
    std::string path, name;
  };
  std::vector<Todo> stack;
  auto emit = [&](const Todo& t, const struct stat& st) {
    if (!taken.insert(t.name).second) {
      warn("skipped (same name already in the archive): " + t.name);
      return;
    }
    Source s;
    s.path = t.path;
    s.item.name = t.name;
    s.item.folder = S_ISDIR(st.st_mode);
    s.item.size = s.item.folder ? 0 : (uint64_t)st.st_size;
    s.item.mtime = (int64_t)st.st_mtime;
    out.push_back(std::move(s));
  };
  auto push_children = [&](const Todo& dir) {
    std::vector<std::string> names;
    if (DIR* d = opendir(dir.path.c_str())) {
      while (const dirent* e = readdir(d)) {
        const std::string n = e->d_name;
        if (n != "." && n != "..") names.push_back(n);
      }
      closedir(d);
    }
    std::sort(names.begin(), names.end());  // (std::string compares bytes as unsigned: the names' byte order)
    for (auto it = names.rbegin(); it != names.rend(); ++it) stack.push_back({dir.path + "/" + *it, dir.name + "/" + *it});
  };
  for (const std::string& in : inputs) {
    struct stat st;
    if (!tree_detail::stat_of(in, st) && stat(in.c_str(), &st) != 0) throw Error("not found: " + in);
    if (S_ISLNK(st.st_mode) && stat(in.c_str(), &st) != 0) throw Error("not found: " + in);  // (an input link is taken)
    Todo top{in, tree_detail::input_name(in)};
    if (!S_ISDIR(st.st_mode) && !S_ISREG(st.st_mode)) {
      warn("skipped (not a file or folder): " + in);
      continue;
    }
    emit(top, st);
    if (!S_ISDIR(st.st_mode)) continue;
    push_children(top);
    while (!stack.empty()) {
      const Todo t = std::move(stack.back());
      stack.pop_back();
      struct stat cs;
      if (!tree_detail::stat_of(t.path, cs) || S_ISLNK(cs.st_mode)) continue;
      if (!S_ISDIR(cs.st_mode) && !S_ISREG(cs.st_mode)) continue;
      emit(t, cs);
      if (S_ISDIR(cs.st_mode)) push_children(t);
    }
  }
  return out;
}

}  // namespace zero
