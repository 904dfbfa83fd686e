// jsonParser.hh -- the subset of CASMcode_global's CASM::jsonParser
// (casm/casm_io/json/jsonParser.hh, CASMcode_global 2.x; not present here)
// that libcasm-monte's JSON readers and writers use: a JSON value with
// object/array construction (put_obj, put_array, operator[], push_back,
// contains), conversion of numbers, strings, booleans, vectors and optionals,
// pretty or compact printing, and reading a document back.  Objects keep their
// members ordered by key, as CASMcode_global's objects do.
#ifndef CASM_casm_io_json_jsonParser_MI355X
#define CASM_casm_io_json_jsonParser_MI355X

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <map>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace CASM {

class jsonParser {
 public:
  enum class Type { null, boolean, integer, number, string, array, object };

  jsonParser() = default;
  jsonParser(std::nullptr_t) {}
  jsonParser(bool v) : m_type(Type::boolean), m_b(v) {}
  template <typename T, typename std::enable_if<std::is_integral<T>::value &&
                                                    !std::is_same<T, bool>::value,
                                                int>::type = 0>
  jsonParser(T v) : m_type(Type::integer), m_i(static_cast<long long>(v)) {}
  jsonParser(double v) : m_type(Type::number), m_d(v) {}
  jsonParser(const char *v) : m_type(Type::string), m_s(v) {}
  jsonParser(std::string v) : m_type(Type::string), m_s(std::move(v)) {}
  template <typename T>
  jsonParser(std::vector<T> const &v) : m_type(Type::array) {
    for (auto const &x : v) m_a.emplace_back(x);
  }
  template <typename T>
  jsonParser(std::optional<T> const &v) {
    if (v.has_value()) *this = jsonParser(*v);
  }

  static jsonParser object() {
    jsonParser j;
    j.m_type = Type::object;
    return j;
  }
  static jsonParser array() {
    jsonParser j;
    j.m_type = Type::array;
    return j;
  }
  static jsonParser null() { return jsonParser(); }

  Type type() const { return m_type; }
  bool is_null() const { return m_type == Type::null; }
  bool is_obj() const { return m_type == Type::object; }
  bool is_array() const { return m_type == Type::array; }
  bool is_string() const { return m_type == Type::string; }
  bool is_bool() const { return m_type == Type::boolean; }
  bool is_number() const { return m_type == Type::number || m_type == Type::integer; }

  jsonParser &put_obj() {
    *this = object();
    return *this;
  }
  jsonParser &put_array() {
    *this = array();
    return *this;
  }
  jsonParser &put_null() {
    *this = jsonParser();
    return *this;
  }

  bool contains(std::string const &key) const { return is_obj() && m_o.count(key) != 0; }

  /// Member access; a null value becomes an object.
  jsonParser &operator[](std::string const &key) {
    if (is_null()) m_type = Type::object;
    if (!is_obj()) throw std::runtime_error("jsonParser: operator[] on a non-object");
    return m_o[key];
  }
  jsonParser const &operator[](std::string const &key) const {
    auto it = m_o.find(key);
    if (!is_obj() || it == m_o.end())
      throw std::runtime_error("jsonParser: no member \"" + key + "\"");
    return it->second;
  }
  jsonParser &operator[](size_t i) { return m_a.at(i); }
  jsonParser const &operator[](size_t i) const { return m_a.at(i); }

  size_t size() const { return is_obj() ? m_o.size() : m_a.size(); }

  /// Append; a null value becomes an array.
  jsonParser &push_back(jsonParser const &v) {
    if (is_null()) m_type = Type::array;
    if (!is_array()) throw std::runtime_error("jsonParser: push_back on a non-array");
    m_a.push_back(v);
    return *this;
  }

  std::map<std::string, jsonParser> const &members() const { return m_o; }
  std::vector<jsonParser> const &elements() const { return m_a; }

  template <typename T>
  T get() const;

  void print(std::ostream &os, int indent = 2) const { _print(os, indent, 0); }

  bool write(std::filesystem::path const &path, int indent = 2) const {
    if (path.has_parent_path()) std::filesystem::create_directories(path.parent_path());
    std::ofstream f(path);
    if (!f) return false;
    print(f, indent);
    f << "\n";
    return static_cast<bool>(f);
  }

  static jsonParser parse(std::string const &text) {
    size_t p = 0;
    jsonParser j = _parse_value(text, p);
    _skip_ws(text, p);
    if (p != text.size()) throw std::runtime_error("jsonParser: trailing characters");
    return j;
  }

  static jsonParser read(std::filesystem::path const &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("jsonParser: cannot read " + path.string());
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
  }

 private:
  Type m_type = Type::null;
  bool m_b = false;
  long long m_i = 0;
  double m_d = 0.0;
  std::string m_s;
  std::vector<jsonParser> m_a;
  std::map<std::string, jsonParser> m_o;

  static void _escape(std::ostream &os, std::string const &s) {
    os << '"';
    for (char c : s) {
      switch (c) {
        case '"': os << "\\\""; break;
        case '\\': os << "\\\\"; break;
        case '\n': os << "\\n"; break;
        case '\t': os << "\\t"; break;
        case '\r': os << "\\r"; break;
        default:
          if (static_cast<unsigned char>(c) < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof(buf), "\\u%04x", c);
            os << buf;
          } else {
            os << c;
          }
      }
    }
    os << '"';
  }

  void _print(std::ostream &os, int indent, int level) const {
    auto nl = [&](int lv) {
      if (indent < 0) return;
      os << '\n' << std::string(static_cast<size_t>(indent * lv), ' ');
    };
    switch (m_type) {
      case Type::null: os << "null"; break;
      case Type::boolean: os << (m_b ? "true" : "false"); break;
      case Type::integer: os << m_i; break;
      case Type::number: {
        if (std::isnan(m_d) || std::isinf(m_d)) {
          os << "null";  // JSON has no NaN/Inf
        } else {
          char buf[32];
          std::snprintf(buf, sizeof(buf), "%.17g", m_d);
          std::string s(buf);
          if (s.find_first_of(".eE") == std::string::npos) s += ".0";
          os << s;
        }
        break;
      }
      case Type::string: _escape(os, m_s); break;
      case Type::array: {
        os << '[';
        for (size_t i = 0; i < m_a.size(); ++i) {
          if (i) os << (indent < 0 ? "," : ",");
          nl(level + 1);
          m_a[i]._print(os, indent, level + 1);
        }
        if (!m_a.empty()) nl(level);
        os << ']';
        break;
      }
      case Type::object: {
        os << '{';
        size_t i = 0;
        for (auto const &p : m_o) {
          if (i++) os << ',';
          nl(level + 1);
          _escape(os, p.first);
          os << (indent < 0 ? ":" : ": ");
          p.second._print(os, indent, level + 1);
        }
        if (!m_o.empty()) nl(level);
        os << '}';
        break;
      }
    }
  }

  static void _skip_ws(std::string const &t, size_t &p) {
    while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\t' || t[p] == '\r')) ++p;
  }

  static std::string _parse_string(std::string const &t, size_t &p) {
    if (t[p] != '"') throw std::runtime_error("jsonParser: expected a string");
    ++p;
    std::string s;
    while (p < t.size() && t[p] != '"') {
      char c = t[p++];
      if (c == '\\') {
        if (p >= t.size()) break;
        char e = t[p++];
        switch (e) {
          case 'n': s += '\n'; break;
          case 't': s += '\t'; break;
          case 'r': s += '\r'; break;
          case 'b': s += '\b'; break;
          case 'f': s += '\f'; break;
          case 'u': {
            unsigned code = std::stoul(t.substr(p, 4), nullptr, 16);
            p += 4;
            if (code < 0x80) {
              s += static_cast<char>(code);
            } else if (code < 0x800) {
              s += static_cast<char>(0xC0 | (code >> 6));
              s += static_cast<char>(0x80 | (code & 0x3F));
            } else {
              s += static_cast<char>(0xE0 | (code >> 12));
              s += static_cast<char>(0x80 | ((code >> 6) & 0x3F));
              s += static_cast<char>(0x80 | (code & 0x3F));
            }
            break;
          }
          default: s += e;
        }
      } else {
        s += c;
      }
    }
    if (p >= t.size()) throw std::runtime_error("jsonParser: unterminated string");
    ++p;
    return s;
  }

  static jsonParser _parse_value(std::string const &t, size_t &p) {
    _skip_ws(t, p);
    if (p >= t.size()) throw std::runtime_error("jsonParser: unexpected end of input");
    char c = t[p];
    if (c == '{') {
      ++p;
      jsonParser j = object();
      _skip_ws(t, p);
      if (t[p] == '}') {
        ++p;
        return j;
      }
      while (true) {
        _skip_ws(t, p);
        std::string key = _parse_string(t, p);
        _skip_ws(t, p);
        if (t[p] != ':') throw std::runtime_error("jsonParser: expected ':'");
        ++p;
        j.m_o[key] = _parse_value(t, p);
        _skip_ws(t, p);
        if (t[p] == ',') {
          ++p;
          continue;
        }
        if (t[p] == '}') {
          ++p;
          return j;
        }
        throw std::runtime_error("jsonParser: expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p;
      jsonParser j = array();
      _skip_ws(t, p);
      if (t[p] == ']') {
        ++p;
        return j;
      }
      while (true) {
        j.m_a.push_back(_parse_value(t, p));
        _skip_ws(t, p);
        if (t[p] == ',') {
          ++p;
          continue;
        }
        if (t[p] == ']') {
          ++p;
          return j;
        }
        throw std::runtime_error("jsonParser: expected ',' or ']'");
      }
    }
    if (c == '"') return jsonParser(_parse_string(t, p));
    if (t.compare(p, 4, "true") == 0) {
      p += 4;
      return jsonParser(true);
    }
    if (t.compare(p, 5, "false") == 0) {
      p += 5;
      return jsonParser(false);
    }
    if (t.compare(p, 4, "null") == 0) {
      p += 4;
      return jsonParser();
    }
    size_t q = p;
    bool is_int = true;
    while (q < t.size() && (std::isdigit(static_cast<unsigned char>(t[q])) || t[q] == '-' ||
                            t[q] == '+' || t[q] == '.' || t[q] == 'e' || t[q] == 'E')) {
      if (t[q] == '.' || t[q] == 'e' || t[q] == 'E') is_int = false;
      ++q;
    }
    if (q == p) throw std::runtime_error("jsonParser: unexpected character");
    std::string num = t.substr(p, q - p);
    p = q;
    if (is_int) return jsonParser(std::stoll(num));
    return jsonParser(std::stod(num));
  }
};

template <>
inline double jsonParser::get<double>() const {
  if (m_type == Type::number) return m_d;
  if (m_type == Type::integer) return static_cast<double>(m_i);
  throw std::runtime_error("jsonParser: not a number");
}
template <>
inline long jsonParser::get<long>() const {
  if (m_type == Type::integer) return static_cast<long>(m_i);
  if (m_type == Type::number && m_d == std::floor(m_d)) return static_cast<long>(m_d);
  throw std::runtime_error("jsonParser: not an integer");
}
template <>
inline int jsonParser::get<int>() const {
  return static_cast<int>(get<long>());
}
template <>
inline bool jsonParser::get<bool>() const {
  if (m_type == Type::boolean) return m_b;
  throw std::runtime_error("jsonParser: not a boolean");
}
template <>
inline std::string jsonParser::get<std::string>() const {
  if (m_type == Type::string) return m_s;
  throw std::runtime_error("jsonParser: not a string");
}
template <>
inline std::vector<double> jsonParser::get<std::vector<double>>() const {
  std::vector<double> v;
  for (auto const &x : m_a) v.push_back(x.get<double>());
  return v;
}
template <>
inline std::vector<long> jsonParser::get<std::vector<long>>() const {
  std::vector<long> v;
  for (auto const &x : m_a) v.push_back(x.get<long>());
  return v;
}
template <>
inline std::vector<std::string> jsonParser::get<std::vector<std::string>>() const {
  std::vector<std::string> v;
  for (auto const &x : m_a) v.push_back(x.get<std::string>());
  return v;
}

}  // namespace CASM

#endif
