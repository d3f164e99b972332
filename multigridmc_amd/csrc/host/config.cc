#include "config.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace mgmc_host {

namespace {

struct Parser {
  const std::string& t;
  std::string origin;
  size_t p = 0;
  int line = 1;

  [[noreturn]] void fail(const std::string& what) const {
    throw ConfigError(origin + ":" + std::to_string(line) + ": " + what);
  }
  char peek() const { return p < t.size() ? t[p] : '\0'; }
  char get() {
    char c = peek();
    if (c == '\n') ++line;
    if (p < t.size()) ++p;
    return c;
  }
  void skip() {
    for (;;) {
      char c = peek();
      if (c == '\0') return;
      if (std::isspace((unsigned char)c)) {
        get();
      } else if (c == '#' || (c == '/' && p + 1 < t.size() && t[p + 1] == '/')) {
        while (peek() != '\n' && peek() != '\0') get();
      } else if (c == '/' && p + 1 < t.size() && t[p + 1] == '*') {
        get();
        get();
        while (!(peek() == '*' && p + 1 < t.size() && t[p + 1] == '/')) {
          if (peek() == '\0') fail("unterminated comment");
          get();
        }
        get();
        get();
      } else {
        return;
      }
    }
  }
  std::string name() {
    std::string n;
    if (!(std::isalpha((unsigned char)peek()) || peek() == '*')) fail("setting name expected");
    while (std::isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-' || peek() == '*') n += get();
    return n;
  }
  std::string string_literal() {
    std::string out;
    for (;;) {
      if (get() != '"') fail("string expected");
      for (;;) {
        char c = get();
        if (c == '\0') fail("unterminated string");
        if (c == '"') break;
        if (c == '\\') {
          char e = get();
          switch (e) {
            case 'n': out += '\n'; break;
            case 't': out += '\t'; break;
            case 'r': out += '\r'; break;
            case 'f': out += '\f'; break;
            case '\\': out += '\\'; break;
            case '"': out += '"'; break;
            case 'x': {
              std::string h;
              h += get();
              h += get();
              out += (char)std::strtol(h.c_str(), nullptr, 16);
              break;
            }
            default: fail("invalid escape in string");
          }
        } else {
          out += c;
        }
      }
      skip();
      if (peek() != '"') return out;  // adjacent literals are concatenated
    }
  }
  Setting scalar() {
    Setting s;
    char c = peek();
    if (c == '"') {
      s.kind = Setting::String;
      s.s = string_literal();
      return s;
    }
    if (std::isalpha((unsigned char)c)) {
      std::string w;
      while (std::isalpha((unsigned char)peek())) w += (char)std::tolower((unsigned char)get());
      if (w == "true" || w == "false") {
        s.kind = Setting::Bool;
        s.b = w == "true";
        return s;
      }
      fail("unknown value '" + w + "'");
    }
    size_t start = p;
    if (c == '+' || c == '-') get();
    bool hex = false, isfloat = false;
    if (peek() == '0' && p + 1 < t.size() && (t[p + 1] == 'x' || t[p + 1] == 'X')) {
      hex = true;
      get();
      get();
      while (std::isxdigit((unsigned char)peek())) get();
    } else {
      while (std::isdigit((unsigned char)peek()) || peek() == '.' || peek() == 'e' || peek() == 'E' ||
             ((peek() == '+' || peek() == '-') && (t[p - 1] == 'e' || t[p - 1] == 'E'))) {
        if (peek() == '.' || peek() == 'e' || peek() == 'E') isfloat = true;
        get();
      }
    }
    std::string num = t.substr(start, p - start);
    if (num.empty() || num == "+" || num == "-") fail("value expected");
    while (peek() == 'L') get();  // 64-bit integer suffix
    char* end = nullptr;
    if (isfloat) {
      s.kind = Setting::Float;
      s.d = std::strtod(num.c_str(), &end);
    } else {
      s.kind = Setting::Int;
      s.i = std::strtoll(num.c_str(), &end, hex ? 16 : 10);
    }
    if (!end || *end != '\0') fail("malformed number '" + num + "'");
    return s;
  }
  Setting value() {
    skip();
    char c = peek();
    Setting s;
    if (c == '{') {
      get();
      s.kind = Setting::Group;
      settings(s, '}');
      get();
    } else if (c == '[' || c == '(') {
      char close = c == '[' ? ']' : ')';
      get();
      s.kind = c == '[' ? Setting::Array : Setting::List;
      skip();
      while (peek() != close) {
        if (c == '[' && (peek() == '{' || peek() == '[' || peek() == '(')) fail("arrays hold scalars only");
        Setting e = c == '[' ? scalar() : value();
        if (c == '[' && e.kind != Setting::Int && e.kind != Setting::Float && e.kind != Setting::Bool &&
            e.kind != Setting::String)
          fail("arrays hold scalars only");
        s.items.push_back(e);
        skip();
        if (peek() == ',') {
          get();
          skip();
        } else if (peek() != close) {
          fail(std::string("',' or '") + close + "' expected");
        }
      }
      get();
    } else {
      s = scalar();
    }
    return s;
  }
  void settings(Setting& group, char close) {
    for (;;) {
      skip();
      if (peek() == close) return;
      if (peek() == '\0') fail(close == '}' ? "'}' expected" : "unexpected end of input");
      std::string n = name();
      skip();
      if (peek() != '=' && peek() != ':') fail("'=' or ':' expected after '" + n + "'");
      get();
      Setting v = value();
      v.name = n;
      for (const auto& m : group.items)
        if (m.name == n) fail("duplicate setting '" + n + "'");
      group.items.push_back(std::move(v));
      skip();
      if (peek() == ';' || peek() == ',') get();
    }
  }
};

const char* kind_name(Setting::Kind k) {
  switch (k) {
    case Setting::Int: return "integer";
    case Setting::Float: return "float";
    case Setting::Bool: return "boolean";
    case Setting::String: return "string";
    case Setting::Array: return "array";
    case Setting::List: return "list";
    default: return "group";
  }
}

}  // namespace

const Setting& Setting::lookup(const std::string& path) const {
  const Setting* cur = this;
  size_t start = 0;
  while (start <= path.size()) {
    size_t dot = path.find('.', start);
    std::string part = path.substr(start, dot == std::string::npos ? std::string::npos : dot - start);
    if (cur->kind != Group) throw ConfigError("setting '" + path + "': '" + cur->name + "' is not a group");
    const Setting* next = nullptr;
    for (const auto& m : cur->items)
      if (m.name == part) next = &m;
    if (!next) throw ConfigError("setting not found: '" + path + "'");
    cur = next;
    if (dot == std::string::npos) break;
    start = dot + 1;
  }
  return *cur;
}

bool Setting::exists(const std::string& path) const {
  try {
    lookup(path);
    return true;
  } catch (const ConfigError&) {
    return false;
  }
}

const Setting& Setting::operator[](size_t idx) const {
  if (kind != Array && kind != List) throw ConfigError("setting '" + name + "' is not an array or list");
  if (idx >= items.size()) throw ConfigError("index out of range in setting '" + name + "'");
  return items[idx];
}

int64_t Setting::as_int() const {
  if (kind != Int) throw ConfigError("setting '" + name + "' is a " + kind_name(kind) + ", integer expected");
  return i;
}

double Setting::as_double() const {
  if (kind == Float) return d;
  if (kind == Int) return (double)i;
  throw ConfigError("setting '" + name + "' is a " + kind_name(kind) + ", number expected");
}

bool Setting::as_bool() const {
  if (kind != Bool) throw ConfigError("setting '" + name + "' is a " + kind_name(kind) + ", boolean expected");
  return b;
}

const std::string& Setting::as_string() const {
  if (kind != String) throw ConfigError("setting '" + name + "' is a " + kind_name(kind) + ", string expected");
  return s;
}

std::vector<double> Setting::as_doubles() const {
  std::vector<double> out;
  for (size_t k = 0; k < length(); ++k) out.push_back((*this)[k].as_double());
  return out;
}

Setting parse_config(const std::string& text, const std::string& origin) {
  Parser ps{text, origin};
  Setting root;
  root.kind = Setting::Group;
  ps.settings(root, '\0');
  return root;
}

Setting read_config_file(const std::string& filename) {
  std::ifstream in(filename);
  if (!in) throw ConfigError("cannot open configuration file '" + filename + "'");
  std::stringstream ss;
  ss << in.rdbuf();
  return parse_config(ss.str(), filename);
}

}  // namespace mgmc_host
