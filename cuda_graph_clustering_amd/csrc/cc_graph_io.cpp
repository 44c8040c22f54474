// cc_graph_io.cpp — the instance files of the command line: the 0/1 matrix text, the "graph" block of a baseline
// result file, SNAP/CSV edge lists and the {"key": ["tag", ...]} files of the tag-similarity graph.
//
// The formats, and what each loader accepts, are the reference's (src/cc_graph.cpp:35-103 and :179-306); the
// reading is this library's. A file is mapped into memory and walked as bytes: the matrix and JSON loaders set the
// cells above the diagonal straight into the bitset rows and mirror them in one pass at the end, the JSON block is
// walked twice (shape, then cells) instead of being copied into per-row vectors, the edge list is parsed in place,
// and the tag file goes through a small tokenizer under a grammar of exactly the shape the dataset files have.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "cc_graph.hpp"
#include "cc_graph_io.hpp"

namespace cc {

// ---- byte access ---------------------------------------------------------------------------------------------

FileBytes::FileBytes(const std::string& path, const char* what) {
  const int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) throw std::runtime_error(std::string(what) + " " + path + " cannot be opened");
  struct stat st {};
  if (::fstat(fd, &st) == 0 && S_ISREG(st.st_mode)) {
    size_ = static_cast<size_t>(st.st_size);
    if (size_ > 0) {
      void* p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd, 0);
      if (p != MAP_FAILED) {
        map_ = p;
        data_ = static_cast<const char*>(p);
      }
    }
  }
  if (!map_) {  // a pipe, a special file or a failed mapping: read it
    size_ = 0;
    char chunk[1 << 16];
    for (ssize_t got; (got = ::read(fd, chunk, sizeof(chunk))) > 0;) copy_.append(chunk, static_cast<size_t>(got));
    data_ = copy_.data();
    size_ = copy_.size();
  }
  ::close(fd);
}

FileBytes::~FileBytes() {
  if (map_) ::munmap(map_, size_);
}

namespace {

// The white space of the "C" locale's isspace.
inline bool IsBlank(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }
inline bool IsDigit(char c) { return c >= '0' && c <= '9'; }

// Appends `value` in decimal.
void PutUnsigned(std::string& out, unsigned long long value) {
  char digits[24];
  int len = 0;
  do {
    digits[len++] = static_cast<char>('0' + value % 10);
    value /= 10;
  } while (value);
  while (len) out.push_back(digits[--len]);
}

// Output through stdio in large blocks; errors are reported once at the end.
class Writer {
 public:
  Writer(const std::string& path, const char* what) : path_(path), what_(what) {
    file_ = std::fopen(path.c_str(), "wb");
    if (!file_) throw std::runtime_error(std::string(what) + " " + path + " cannot be written");
  }
  ~Writer() {
    if (file_) std::fclose(file_);
  }
  std::string& buffer() { return buffer_; }
  void Flush(bool force = false) {
    if (buffer_.size() >= (1u << 20) || force) {
      ok_ &= std::fwrite(buffer_.data(), 1, buffer_.size(), file_) == buffer_.size();
      buffer_.clear();
    }
  }
  void Close() {
    Flush(true);
    ok_ &= std::fclose(file_) == 0;
    file_ = nullptr;
    if (!ok_) throw std::runtime_error(std::string(what_) + " " + path_ + ": write failed");
  }

 private:
  std::string path_;
  const char* what_;
  std::FILE* file_ = nullptr;
  std::string buffer_;
  bool ok_ = true;
};

// Graph density as an ostream prints a double by default (six significant digits, %g).
std::string DensityText(double d) {
  char text[32];
  std::snprintf(text, sizeof(text), "%g", d);
  return text;
}

}  // namespace

// ---- the bitset from its upper triangle ----------------------------------------------------------------------

void Graph::MirrorUpperTriangle() {
  uint64_t edges = 0;
  for (unsigned i = 0; i < n_; ++i) {
    const Word* row = bits_.data() + static_cast<size_t>(i) * words_per_row_;
    for (unsigned w = (i + 1) / kWordBits; w < words_per_row_; ++w) {
      Word above = row[w];
      if (w == (i + 1) / kWordBits) above &= ~Word(0) << ((i + 1) % kWordBits);  // only j > i
      for (; above; above &= above - 1) {
        const unsigned j = w * kWordBits + static_cast<unsigned>(__builtin_ctzll(above));
        bits_[static_cast<size_t>(j) * words_per_row_ + i / kWordBits] |= Word(1) << (i % kWordBits);
        ++edges;
      }
    }
  }
  edges_ = edges;
}

// ---- 0/1 matrix text -----------------------------------------------------------------------------------------

// "n", then n x n cells, each the next byte that is not white space ('1': joined, anything else: not), so both
// "0 1 0" rows and "010" rows read. Only the cells above the diagonal decide; the matrix is made symmetric from them
// and the diagonal stays empty (reference src/cc_graph.cpp:179-198).
Graph Graph::LoadMatrix(const std::string& path) {
  const FileBytes file(path, "graph file");
  const char* p = file.begin();
  const char* const end = file.end();
  while (p != end && IsBlank(*p)) ++p;
  if (p != end && *p == '+') ++p;
  uint64_t n = 0;
  const char* first_digit = p;
  while (p != end && IsDigit(*p) && n <= 0xFFFFFFFFull) n = n * 10 + static_cast<uint64_t>(*p++ - '0');
  if (p == first_digit || n == 0 || n > 0xFFFFFFFFull) {
    throw std::runtime_error("graph file " + path + ": the first entry must be the vertex count n >= 1");
  }
  Graph g(static_cast<unsigned>(n));
  const unsigned size = g.n_;
  for (unsigned i = 0; i < size; ++i) {
    Word* row = g.bits_.data() + static_cast<size_t>(i) * g.words_per_row_;
    for (unsigned j = 0; j < size; ++j) {
      while (p != end && IsBlank(*p)) ++p;
      if (p == end) {
        throw std::runtime_error("graph file " + path + " ends inside the matrix (row " + std::to_string(i) +
                                 ", column " + std::to_string(j) + ")");
      }
      if (*p++ == '1' && j > i) row[j / kWordBits] |= Word(1) << (j % kWordBits);
    }
  }
  g.MirrorUpperTriangle();
  return g;
}

void Graph::SaveMatrix(const std::string& path) const {
  Writer out(path, "graph file");
  std::string& b = out.buffer();
  PutUnsigned(b, n_);
  b.push_back('\n');
  for (unsigned i = 0; i < n_; ++i) {
    const Word* row = Row(i);
    for (unsigned j = 0; j < n_; ++j) b.push_back(static_cast<char>('0' + ((row[j / kWordBits] >> (j % kWordBits)) & 1)));
    b.push_back('\n');
    out.Flush();
  }
  out.Close();
}

// ---- the "graph" block of a baseline result file -------------------------------------------------------------

namespace {

// The bracketed block that starts at the first '[' after the key "graph": its lists at depth 2 are the rows, the
// bytes '0' and '1' inside them the cells; everything else, deeper lists included, is skipped, and the walk ends
// where the block closes or the file ends (reference src/cc_graph.cpp:213-252). Calls row_end(cells) for every row
// and cell(row, column, value) for every cell.
template <typename RowEnd, typename Cell>
void WalkGraphBlock(const char* p, const char* end, RowEnd&& row_end, Cell&& cell) {
  int depth = 0;
  unsigned row = 0, column = 0;
  for (; p != end; ++p) {
    const char c = *p;
    if (c == '[') {
      if (++depth == 2) column = 0;
    } else if (c == ']') {
      if (depth == 2) row_end(row++, column);
      if (--depth == 0) return;
    } else if (depth == 2 && (c == '0' || c == '1')) {
      cell(row, column++, c == '1');
    }
  }
}

}  // namespace

Graph Graph::LoadBaselineJson(const std::string& path) {
  const FileBytes file(path, "baseline json");
  const char* const end = file.end();
  static const char kKey[] = "\"graph\"";
  const char* key = std::search(file.begin(), end, kKey, kKey + sizeof(kKey) - 1);
  if (key == end) throw std::runtime_error("baseline json " + path + " has no \"graph\" key");
  const char* open = std::find(key, end, '[');
  if (open == end) throw std::runtime_error("baseline json " + path + ": no matrix follows the \"graph\" key");
  // first walk: the shape (n = the number of rows; every row must hold n cells)
  std::vector<unsigned> row_cells;
  WalkGraphBlock(open, end, [&](unsigned, unsigned cells) { row_cells.push_back(cells); },
                 [](unsigned, unsigned, bool) {});
  if (row_cells.empty()) throw std::runtime_error("baseline json " + path + ": the \"graph\" matrix has no rows");
  const unsigned n = static_cast<unsigned>(row_cells.size());
  for (unsigned i = 0; i < n; ++i) {
    if (row_cells[i] != n) {
      throw std::runtime_error("baseline json " + path + ": \"graph\" row " + std::to_string(i) + " has " +
                               std::to_string(row_cells[i]) + " cells, not " + std::to_string(n));
    }
  }
  // second walk: the cells above the diagonal into the rows
  Graph g(n);
  WalkGraphBlock(open, end, [](unsigned, unsigned) {}, [&](unsigned i, unsigned j, bool one) {
    if (one && j > i) g.bits_[static_cast<size_t>(i) * g.words_per_row_ + j / kWordBits] |= Word(1) << (j % kWordBits);
  });
  g.MirrorUpperTriangle();
  return g;
}

// Byte for byte what the reference writes (src/cc_graph.cpp:254-268), so its files and ours are interchangeable.
void Graph::SaveBaselineJson(const std::string& path) const {
  Writer out(path, "baseline json");
  std::string& b = out.buffer();
  b += "{\n\"size\": ";
  PutUnsigned(b, n_);
  b += ",\n\"density\": " + DensityText(Density()) + ",\n\"graph\": [\n";
  for (unsigned i = 0; i < n_; ++i) {
    const Word* row = Row(i);
    b.push_back('[');
    for (unsigned j = 0; j < n_; ++j) {
      if (j) b.push_back(',');
      b.push_back(static_cast<char>('0' + ((row[j / kWordBits] >> (j % kWordBits)) & 1)));
    }
    b += i + 1 < n_ ? "],\n" : "]\n";
    out.Flush();
  }
  b += "]\n}\n";
  out.Close();
}

// ---- edge lists ----------------------------------------------------------------------------------------------

// One edge per line: two decimal vertex ids separated (and preceded) by spaces, tabs or commas; the rest of the line
// (weights, times) is not read. Lines that start with '#' or '%', and lines that do not hold two ids there (a CSV
// header, a blank line), are skipped. A vertex is the rank of its id among all ids of the file; direction, loops and
// repeats are dropped (reference src/cc_graph.cpp:270-306).
Graph Graph::LoadEdgeList(const std::string& path) {
  const FileBytes file(path, "edge list");
  const char* p = file.begin();
  const char* const end = file.end();
  std::vector<uint64_t> ends;
  auto read_id = [&](uint64_t* id) {
    while (p != end && (*p == ' ' || *p == '\t' || *p == ',')) ++p;
    if (p == end || !IsDigit(*p)) return false;
    uint64_t value = 0;
    for (; p != end && IsDigit(*p); ++p) {
      const uint64_t digit = static_cast<uint64_t>(*p - '0');
      if (value > (UINT64_MAX - digit) / 10) {
        throw std::runtime_error("edge list " + path + ": a vertex id exceeds 64 bits at byte " +
                                 std::to_string(p - file.begin()));
      }
      value = value * 10 + digit;
    }
    *id = value;
    return true;
  };
  while (p != end) {
    const char* line_end = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    if (!line_end) line_end = end;
    uint64_t u = 0, v = 0;
    if (*p != '#' && *p != '%' && read_id(&u) && read_id(&v)) {
      ends.push_back(u);
      ends.push_back(v);
    }
    p = line_end == end ? end : line_end + 1;
  }
  if (ends.empty()) throw std::runtime_error("edge list " + path + " holds no edge");
  std::vector<uint64_t> ids(ends);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  if (ids.size() > 0xFFFFFFFFull) throw std::runtime_error("edge list " + path + " names more than 2^32 - 1 vertices");
  Graph g(static_cast<unsigned>(ids.size()));
  for (size_t e = 0; e < ends.size(); e += 2) {
    const unsigned a = static_cast<unsigned>(std::lower_bound(ids.begin(), ids.end(), ends[e]) - ids.begin());
    const unsigned b = static_cast<unsigned>(std::lower_bound(ids.begin(), ids.end(), ends[e + 1]) - ids.begin());
    if (a == b) continue;
    const unsigned lo = std::min(a, b), hi = std::max(a, b);
    g.bits_[static_cast<size_t>(lo) * g.words_per_row_ + hi / kWordBits] |= Word(1) << (hi % kWordBits);
  }
  g.MirrorUpperTriangle();
  return g;
}

void Graph::SaveEdgeList(const std::string& path) const {
  Writer out(path, "edge list");
  std::string& b = out.buffer();
  b += "# ";
  PutUnsigned(b, n_);
  b += " vertices, ";
  PutUnsigned(b, edges_);
  b += " edges\n";
  for (unsigned i = 0; i < n_; ++i) {
    const Word* row = Row(i);
    for (unsigned w = (i + 1) / kWordBits; w < words_per_row_; ++w) {
      Word above = row[w];
      if (w == (i + 1) / kWordBits) above &= ~Word(0) << ((i + 1) % kWordBits);
      for (; above; above &= above - 1) {
        PutUnsigned(b, i);
        b.push_back(' ');
        PutUnsigned(b, w * kWordBits + static_cast<unsigned>(__builtin_ctzll(above)));
        b.push_back('\n');
      }
    }
    out.Flush();
  }
  out.Close();
}

// ---- tag files -------------------------------------------------------------------------------------------------

namespace {

// The tokens of a tag file.
enum class Tok { kOpenObject, kCloseObject, kOpenList, kCloseList, kColon, kComma, kString, kEnd, kOther };

class TagLexer {
 public:
  TagLexer(const FileBytes& file, const std::string& path) : p_(file.begin()), begin_(file.begin()), end_(file.end()),
                                                              path_(path) {}

  // The next token; a string's text (a backslash takes the next byte as it is) is left in text().
  Tok Next() {
    while (p_ != end_ && IsBlank(*p_)) ++p_;
    at_ = p_;
    if (p_ == end_) return Tok::kEnd;
    switch (*p_++) {
      case '{': return Tok::kOpenObject;
      case '}': return Tok::kCloseObject;
      case '[': return Tok::kOpenList;
      case ']': return Tok::kCloseList;
      case ':': return Tok::kColon;
      case ',': return Tok::kComma;
      case '"': break;
      default: return Tok::kOther;
    }
    text_.clear();
    while (true) {
      if (p_ == end_) Fail("a string is not closed");
      const char c = *p_++;
      if (c == '"') return Tok::kString;
      if (c == '\\') {
        if (p_ == end_) Fail("a string is not closed");
        text_.push_back(*p_++);
      } else {
        text_.push_back(c);
      }
    }
  }

  void Expect(Tok want, const char* what) {
    if (Next() != want) Fail(what);
  }

  const std::string& text() const { return text_; }

  [[noreturn]] void Fail(const char* what) const {
    throw std::runtime_error("tag file " + path_ + ": " + what + " at byte " + std::to_string(at_ - begin_));
  }

 private:
  const char* p_;
  const char* begin_;
  const char* end_;
  const char* at_ = nullptr;
  const std::string& path_;
  std::string text_;
};

}  // namespace

// {"key": ["tag", ...], ...}: per object, in file order, its distinct tags as ids in increasing order. Keys are not
// read (an object's vertex is its position); after the closing brace the file is not read either.
std::vector<std::vector<int>> ReadTagFile(const std::string& path) {
  const FileBytes file(path, "tag file");
  TagLexer lex(file, path);
  std::unordered_map<std::string, int> tag_id;
  std::vector<std::vector<int>> objects;
  lex.Expect(Tok::kOpenObject, "expected '{'");
  Tok t = lex.Next();
  if (t == Tok::kCloseObject) return objects;
  while (true) {
    if (t != Tok::kString) lex.Fail("expected an object key");
    lex.Expect(Tok::kColon, "expected ':'");
    lex.Expect(Tok::kOpenList, "expected '['");
    std::vector<int> tags;
    t = lex.Next();
    if (t != Tok::kCloseList) {
      while (true) {
        if (t != Tok::kString) lex.Fail("expected a tag string");
        tags.push_back(tag_id.try_emplace(lex.text(), static_cast<int>(tag_id.size())).first->second);
        t = lex.Next();
        if (t == Tok::kCloseList) break;
        if (t != Tok::kComma) lex.Fail("expected ',' or ']'");
        t = lex.Next();
      }
    }
    std::sort(tags.begin(), tags.end());
    tags.erase(std::unique(tags.begin(), tags.end()), tags.end());
    objects.push_back(std::move(tags));
    t = lex.Next();
    if (t == Tok::kCloseObject) return objects;
    if (t != Tok::kComma) lex.Fail("expected ',' or '}'");
    t = lex.Next();
  }
}

}  // namespace cc
