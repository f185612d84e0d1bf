// Config.hpp -- the configuration tree of the reference and its "mf" text
// format, for driving the C ABI from the reference's own input files.
//
// Tree: Config/Config.hpp:28 (a Boost iptree: ordered children, keys
// compared without regard to case, values kept as written, '.'-separated
// paths).  Reader: Config/mf_parser/mf_parser_read.hpp:185-402 with the
// settings of mf_parser_config.hpp (comment '!', key/value delimiter "==",
// multi-word keys joined by one space).  The grammar, line by line:
//   #include "file"      another file read into the current block (a relative
//                        path is taken from the including file's directory)
//   words ... == data    a child whose key is the words before "==" and whose
//                        value is the rest of the line (or a quoted string,
//                        which may continue on the next line after a '\')
//   words ...            a child with an empty value (no "==" on the line)
//   {  ...  }            the children of the entry just read
// A '{' or '}' counts only where a key or a value would start; elsewhere it
// is part of the key or value text.  Backslash escapes (\n, \t, \", \\, ...)
// are expanded in keys and values.  Errors carry "file(line): message".
#pragma once

#include <algorithm>
#include <cctype>
#include <fstream>
#include <istream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mfsv {

// ptree_bad_path / ptree_bad_data / file_parser_error of the reference.
struct ConfigBadPath : std::runtime_error {
  explicit ConfigBadPath(const std::string& p) : std::runtime_error("No such node (" + p + ")") {}
};
struct ConfigBadData : std::runtime_error {
  explicit ConfigBadData(const std::string& what) : std::runtime_error(what) {}
};
struct ConfigParseError : std::runtime_error {
  ConfigParseError(const std::string& msg, const std::string& file, unsigned long line)
      : std::runtime_error((file.empty() ? std::string("<unspecified file>") : file) + "(" +
                           std::to_string(line) + "): " + msg),
        message(msg), line_no(line) {}
  std::string message;
  unsigned long line_no;
};

namespace config_detail {

inline std::string lower(std::string s) {
  for (char& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
  return s;
}
inline bool same_key(const std::string& a, const std::string& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::tolower(static_cast<unsigned char>(a[i])) != std::tolower(static_cast<unsigned char>(b[i])))
      return false;
  return true;
}

// Value conversion as the tree's stream translator: the text is read by an
// input stream, trailing blanks are allowed, anything else left over fails.
template <typename T>
bool convert(const std::string& s, T* out) {
  std::istringstream is(s);
  T v{};
  is >> v;
  if (is.fail() || is.bad()) return false;
  is >> std::ws;
  if (!is.eof()) return false;
  *out = v;
  return true;
}
template <>
inline bool convert<std::string>(const std::string& s, std::string* out) {
  *out = s;
  return true;
}
template <>
inline bool convert<bool>(const std::string& s, bool* out) {
  {  // "true" / "false" first, then 0 / 1
    std::istringstream is(s);
    bool v = false;
    is >> std::boolalpha >> v;
    if (!is.fail()) {
      is >> std::ws;
      if (is.eof()) { *out = v; return true; }
    }
  }
  std::istringstream is(s);
  bool v = false;
  is >> v;
  if (is.fail()) return false;
  is >> std::ws;
  if (!is.eof()) return false;
  *out = v;
  return true;
}

}  // namespace config_detail

class Config {
 public:
  using Entry = std::pair<std::string, Config>;

  Config() = default;
  explicit Config(std::string data) : data_(std::move(data)) {}

  const std::string& data() const { return data_; }
  void set_data(std::string d) { data_ = std::move(d); }
  bool empty() const { return children_.empty(); }
  std::vector<Entry>::const_iterator begin() const { return children_.begin(); }
  std::vector<Entry>::const_iterator end() const { return children_.end(); }
  size_t size() const { return children_.size(); }

  // Children whose key equals `key` (no path), in file order.
  std::vector<const Config*> equal_range(const std::string& key) const {
    std::vector<const Config*> out;
    for (const auto& e : children_)
      if (config_detail::same_key(e.first, key)) out.push_back(&e.second);
    return out;
  }
  size_t count(const std::string& key) const { return equal_range(key).size(); }

  // First node on a '.'-separated path, or null.
  const Config* find(const std::string& path) const {
    const Config* node = this;
    size_t a = 0;
    while (node) {
      const size_t b = path.find('.', a);
      const std::string k = path.substr(a, b == std::string::npos ? std::string::npos : b - a);
      const Config* next = nullptr;
      for (const auto& e : node->children_)
        if (config_detail::same_key(e.first, k)) { next = &e.second; break; }
      node = next;
      if (b == std::string::npos) break;
      a = b + 1;
    }
    return node;
  }
  const Config& get_child(const std::string& path) const {
    const Config* c = find(path);
    if (!c) throw ConfigBadPath(path);
    return *c;
  }
  const Config& get_child(const std::string& path, const Config& def) const {
    const Config* c = find(path);
    return c ? *c : def;
  }

  template <typename T>
  T get_value() const {
    T v{};
    if (!config_detail::convert(data_, &v))
      throw ConfigBadData("conversion of data \"" + data_ + "\" failed");
    return v;
  }
  template <typename T>
  T get_value(const T& def) const {
    T v{};
    return config_detail::convert(data_, &v) ? v : def;
  }
  std::string get_value(const char* def) const { return get_value<std::string>(std::string(def)); }

  template <typename T>
  T get(const std::string& path) const {
    const Config& c = get_child(path);
    T v{};
    if (!config_detail::convert(c.data_, &v))
      throw ConfigBadData("conversion of data \"" + c.data_ + "\" at " + path + " failed");
    return v;
  }
  // A missing node or a value that does not convert gives the default (the
  // tree's get with a default value).
  template <typename T>
  T get(const std::string& path, const T& def) const {
    const Config* c = find(path);
    if (!c) return def;
    T v{};
    return config_detail::convert(c->data_, &v) ? v : def;
  }
  std::string get(const std::string& path, const char* def) const {
    return get<std::string>(path, std::string(def));
  }

  // New child at the end (key without path).
  Config& add(const std::string& key, const std::string& value = std::string()) {
    children_.emplace_back(key, Config(value));
    return children_.back().second;
  }
  // Set the value at a path, creating the nodes that are missing.
  Config& put(const std::string& path, const std::string& value) {
    Config* node = this;
    size_t a = 0;
    while (true) {
      const size_t b = path.find('.', a);
      const std::string k = path.substr(a, b == std::string::npos ? std::string::npos : b - a);
      Config* next = nullptr;
      for (auto& e : node->children_)
        if (config_detail::same_key(e.first, k)) { next = &e.second; break; }
      if (!next) next = &node->add(k);
      node = next;
      if (b == std::string::npos) break;
      a = b + 1;
    }
    node->data_ = value;
    return *node;
  }

 private:
  friend class MfReader;
  std::string data_;
  std::vector<Entry> children_;
};

// The mf reader (mf_parser_read.hpp:185-402).
class MfReader {
 public:
  static Config read_file(const std::string& filename) {
    std::ifstream f(filename);
    if (!f) throw ConfigParseError("cannot open file for reading", filename, 0);
    Config root;
    read(f, root, filename, 0);
    return root;
  }
  static Config read_string(const std::string& text, const std::string& filename = std::string()) {
    std::istringstream f(text);
    Config root;
    read(f, root, filename, 0);
    return root;
  }

 private:
  static constexpr char kComment = '!';

  static bool space(char c) {
    const unsigned n = static_cast<unsigned char>(c);
    return n <= 127 && std::isspace(static_cast<int>(n)) != 0;
  }
  static void skip_ws(const char*& t) {
    while (space(*t)) ++t;
  }
  static std::string unescape(const char* b, const char* e) {
    std::string r;
    while (b != e) {
      if (*b == '\\') {
        ++b;
        if (b == e) throw ConfigParseError("character expected after backslash", "", 0);
        switch (*b) {
          case '0': r += '\0'; break;
          case 'a': r += '\a'; break;
          case 'b': r += '\b'; break;
          case 'f': r += '\f'; break;
          case 'n': r += '\n'; break;
          case 'r': r += '\r'; break;
          case 't': r += '\t'; break;
          case 'v': r += '\v'; break;
          case '"': r += '"'; break;
          case '\'': r += '\''; break;
          case '\\': r += '\\'; break;
          default: throw ConfigParseError("unknown escape sequence", "", 0);
        }
      } else {
        r += *b;
      }
      ++b;
    }
    return r;
  }
  static std::string word(const char*& t) {
    skip_ws(t);
    const char* s = t;
    while (!space(*t) && *t != kComment && *t != '\0') ++t;
    return unescape(s, t);
  }
  static std::string rest_of_line(const char*& t) {
    skip_ws(t);
    const char* s = t;
    while (*t != '\0' && *t != kComment) ++t;
    const char* e = t;
    while (e > s && space(*(e - 1))) --e;
    return unescape(s, e);
  }
  // "..." with escapes; *more (when given) says whether a trailing '\'
  // continues the value on the next line.
  static std::string quoted(const char*& t, bool* more) {
    skip_ws(t);
    if (*t != '"') throw ConfigParseError("expected \"", "", 0);
    ++t;
    bool esc = false;
    const char* s = t;
    while ((esc || *t != '"') && *t != '\0') {
      esc = !esc && *t == '\\';
      ++t;
    }
    if (*t != '"') throw ConfigParseError("unexpected end of line", "", 0);
    std::string r = unescape(s, t++);
    skip_ws(t);
    if (*t == '\\') {
      if (!more) throw ConfigParseError("unexpected \\", "", 0);
      ++t;
      skip_ws(t);
      if (*t == '\0' || *t == kComment) *more = true;
      else throw ConfigParseError("expected end of line after \\", "", 0);
    } else if (more) {
      *more = false;
    }
    return r;
  }
  static std::string key(const char*& t) {
    skip_ws(t);
    if (*t == '"') return quoted(t, nullptr);
    std::string k;
    for (std::string w = word(t); !w.empty() && w != "=="; w = word(t)) k += w + " ";
    if (!k.empty()) k.pop_back();
    return k;
  }
  static std::string value(const char*& t, bool* more) {
    skip_ws(t);
    if (*t == '"') return quoted(t, more);
    *more = false;
    return rest_of_line(t);
  }

  static void read(std::istream& in, Config& root, const std::string& filename, int depth) {
    enum State { kKey, kData, kDataCont } state = kKey;
    unsigned long line_no = 0;
    Config* last = nullptr;
    std::vector<Config*> stack{&root};
    std::string line;
    try {
      while (in.good()) {
        ++line_no;
        std::getline(in, line);
        if (!in.good() && !in.eof()) throw ConfigParseError("read error", filename, line_no);
        const char* t = line.c_str();
        skip_ws(t);
        if (*t == '#') {
          ++t;
          const std::string directive = word(t);
          if (directive != "include") throw ConfigParseError("unknown directive", filename, line_no);
          if (depth > 100)
            throw ConfigParseError("include depth too large, probably recursive include", filename,
                                   line_no);
          const std::string inc = quoted(t, nullptr);
          std::string path = inc;
          if (!inc.empty() && inc[0] != '/') {
            const size_t slash = filename.find_last_of('/');
            if (slash != std::string::npos) path = filename.substr(0, slash + 1) + inc;
          }
          std::ifstream f(path);
          if (!f.good()) throw ConfigParseError("cannot open include file " + inc, filename, line_no);
          read(f, *stack.back(), inc, depth + 1);
          skip_ws(t);
          if (*t != '\0') throw ConfigParseError("expected end of line", filename, line_no);
          continue;
        }
        while (true) {
          skip_ws(t);
          if (*t == '\0' || *t == kComment) {
            if (state == kData) state = kKey;
            break;
          }
          switch (state) {
            case kKey:
              if (*t == '{') {
                if (!last) throw ConfigParseError("unexpected {", "", 0);
                stack.push_back(last);
                last = nullptr;
                ++t;
              } else if (*t == '}') {
                if (stack.size() <= 1) throw ConfigParseError("unmatched }", "", 0);
                stack.pop_back();
                last = nullptr;
                ++t;
              } else {
                const std::string k = key(t);
                last = &stack.back()->add(k);
                state = kData;
              }
              break;
            case kData:
              if (*t == '{') {
                stack.push_back(last);
                last = nullptr;
                ++t;
                state = kKey;
              } else if (*t == '}') {
                if (stack.size() <= 1) throw ConfigParseError("unmatched }", "", 0);
                stack.pop_back();
                last = nullptr;
                ++t;
                state = kKey;
              } else {
                bool more = false;
                last->data_ = value(t, &more);
                state = more ? kDataCont : kKey;
              }
              break;
            case kDataCont:
              if (*t == '"') {
                bool more = false;
                last->data_ += quoted(t, &more);
                state = more ? kDataCont : kKey;
              } else {
                throw ConfigParseError("expected \" after \\ in previous line", "", 0);
              }
              break;
          }
        }
      }
      if (stack.size() != 1) throw ConfigParseError("unmatched {", "", 0);
    } catch (const ConfigParseError& e) {
      if (e.line_no == 0) throw ConfigParseError(e.message, filename, line_no);
      throw;
    }
  }
};

// split_string of Config.cpp:22-72: items between any of the delimiters.
// The string form trims each item; the numeric form converts each item as
// written (no blanks allowed around a number).
inline std::vector<std::string> split_strings(const std::string& in,
                                              const std::string& delims = ",") {
  std::vector<std::string> out;
  size_t a = 0;
  while (true) {
    const size_t b = in.find_first_of(delims, a);
    std::string item = in.substr(a, b == std::string::npos ? std::string::npos : b - a);
    const size_t f = item.find_first_not_of(" \t\r\n\v\f");
    const size_t l = item.find_last_not_of(" \t\r\n\v\f");
    out.push_back(f == std::string::npos ? std::string() : item.substr(f, l - f + 1));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return out;
}

namespace config_detail {
// lexical conversion: the whole text must be the number, no blanks
template <typename T>
T lexical(const std::string& s) {
  if (s.empty() || std::isspace(static_cast<unsigned char>(s.front())) ||
      std::isspace(static_cast<unsigned char>(s.back())))
    throw ConfigBadData("bad lexical cast of \"" + s + "\"");
  std::istringstream is(s);
  T v{};
  is >> v;
  if (is.fail() || is.peek() != std::char_traits<char>::eof())
    throw ConfigBadData("bad lexical cast of \"" + s + "\"");
  if constexpr (std::is_unsigned<T>::value)
    if (s.front() == '-') throw ConfigBadData("bad lexical cast of \"" + s + "\"");
  return v;
}
}  // namespace config_detail

template <typename T>
std::vector<T> split_values(const std::string& in, const std::string& delims = ",") {
  std::vector<T> out;
  size_t a = 0;
  while (true) {
    const size_t b = in.find_first_of(delims, a);
    out.push_back(config_detail::lexical<T>(
        in.substr(a, b == std::string::npos ? std::string::npos : b - a)));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return out;
}

// Fixed-length form: trimmed items, exactly D of them.
template <typename T, size_t D>
std::vector<T> split_fixed(const std::string& in, const std::string& delims = ",") {
  const std::vector<std::string> items = split_strings(in, delims);
  if (items.size() != D)
    throw std::runtime_error("Unexpected number of components in vector \"" + in + "\" (expected " +
                             std::to_string(D) + ")");
  std::vector<T> out;
  for (const auto& s : items) out.push_back(config_detail::lexical<T>(s));
  return out;
}

}  // namespace mfsv
