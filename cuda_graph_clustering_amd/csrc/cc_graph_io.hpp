// cc_graph_io.hpp — byte access to input files and the tag-file reader, shared by the loaders (cc_graph_io.cpp) and
// the tag-similarity graph builder (cc_host.cpp).
#pragma once

#include <cstddef>
#include <string>
#include <vector>

namespace cc {

// The whole content of a file, read-only: mapped for a regular file, read into memory otherwise. Throws
// std::runtime_error naming `what` and the path when the file cannot be opened.
class FileBytes {
 public:
  FileBytes(const std::string& path, const char* what);
  ~FileBytes();
  FileBytes(const FileBytes&) = delete;
  FileBytes& operator=(const FileBytes&) = delete;
  const char* begin() const { return data_; }
  const char* end() const { return data_ + size_; }
  size_t size() const { return size_; }

 private:
  void* map_ = nullptr;
  const char* data_ = "";
  size_t size_ = 0;
  std::string copy_;
};

// The objects of a {"key": ["tag", ...], ...} file in file order, each as its distinct tag ids in increasing order
// (ids number the distinct tag strings of the file). Throws std::runtime_error on a file of any other shape.
std::vector<std::vector<int>> ReadTagFile(const std::string& path);

}  // namespace cc
