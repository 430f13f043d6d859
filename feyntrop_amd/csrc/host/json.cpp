/*
 * json.cpp -- the minimal JSON reader/writer the CLI contract needs.
 *
 * Output conventions follow the reference's writer (nlohmann/json 3.11.2 as
 * used by srcs/feyntrop.cpp:271-278): object keys in sorted order, doubles in
 * shortest round-trip digits laid out like its format_buffer (decimal
 * notation for exponents in (-4, 15], otherwise d.ddde+XX with at least two
 * exponent digits, ".0" appended to integral values), NaN/Inf as null.
 */
#include <charconv>
#include <cmath>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <string>

#include "ftr_host.hpp"

namespace ftrh {

const Json* Json::get(const std::string& k) const {
  if (t != Obj) return nullptr;
  for (auto& kv : o)
    if (kv.first == k) return &kv.second;
  return nullptr;
}

uint64_t Json::as_u64() const {
  if (is_int && !is_neg) return u;
  if (is_int && is_neg) return (uint64_t)(-(int64_t)u);
  return (uint64_t)num;
}

int64_t Json::as_i64() const {
  if (is_int) return is_neg ? -(int64_t)u : (int64_t)u;
  return (int64_t)num;
}

namespace {
struct Parser {
  const std::string& s;
  size_t i = 0;
  explicit Parser(const std::string& t) : s(t) {}
  [[noreturn]] void fail(const char* what) {
    throw Error(-1, std::string("[json.exception.parse_error] ") + what + " at byte " + std::to_string(i));
  }
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
  }
  Json value() {
    ws();
    if (i >= s.size()) fail("unexpected end of input");
    const char c = s[i];
    Json v;
    if (c == '{') {
      v.t = Json::Obj;
      i++;
      ws();
      if (i < s.size() && s[i] == '}') { i++; return v; }
      while (true) {
        ws();
        if (i >= s.size() || s[i] != '"') fail("expected string key");
        std::string k = str();
        ws();
        if (i >= s.size() || s[i] != ':') fail("expected ':'");
        i++;
        Json x = value();
        bool found = false;
        for (auto& kv : v.o)
          if (kv.first == k) { kv.second = x; found = true; }
        if (!found) v.o.emplace_back(k, x);
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == '}') { i++; break; }
        fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.t = Json::Arr;
      i++;
      ws();
      if (i < s.size() && s[i] == ']') { i++; return v; }
      while (true) {
        v.a.push_back(value());
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == ']') { i++; break; }
        fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.t = Json::Str;
      v.s = str();
    } else if (s.compare(i, 4, "true") == 0) {
      v.t = Json::Bool; v.b = true; i += 4;
    } else if (s.compare(i, 5, "false") == 0) {
      v.t = Json::Bool; v.b = false; i += 5;
    } else if (s.compare(i, 4, "null") == 0) {
      v.t = Json::Null; i += 4;
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      const size_t b = i;
      if (s[i] == '-') i++;
      bool integer = true;
      while (i < s.size() && ((s[i] >= '0' && s[i] <= '9') || s[i] == '.' || s[i] == 'e' || s[i] == 'E' ||
                              s[i] == '+' || s[i] == '-')) {
        if (s[i] == '.' || s[i] == 'e' || s[i] == 'E') integer = false;
        i++;
      }
      const std::string tok = s.substr(b, i - b);
      v.t = Json::Num;
      char* end = nullptr;
      v.num = std::strtod(tok.c_str(), &end);
      if (!end || *end) fail("invalid number");
      if (integer) {
        v.is_neg = tok[0] == '-';
        errno = 0;
        v.u = std::strtoull(tok.c_str() + (v.is_neg ? 1 : 0), &end, 10);
        v.is_int = (errno == 0);
      }
    } else {
      fail("syntax error");
    }
    return v;
  }
  std::string str() {
    std::string out;
    i++;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\' && i + 1 < s.size()) {
        i++;
        const char e = s[i];
        if (e == 'n') out += '\n';
        else if (e == 't') out += '\t';
        else if (e == 'r') out += '\r';
        else if (e == 'b') out += '\b';
        else if (e == 'f') out += '\f';
        else if (e == 'u') { out += '?'; i += 4; }
        else out += e;
        i++;
      } else {
        out += s[i++];
      }
    }
    if (i >= s.size()) fail("unterminated string");
    i++;
    return out;
  }
};
}  // namespace

Json json_parse(const std::string& text) {
  Parser p(text);
  Json v = p.value();
  p.ws();
  if (p.i != text.size()) p.fail("unexpected trailing input");
  return v;
}

std::string json_number(double v) {
  if (!std::isfinite(v)) return "null";
  if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  *r.ptr = 0;
  /* buf: [-]d[.ddd]e[+-]XX -> digits + decimal exponent */
  std::string s(buf), out;
  if (s[0] == '-') { out = "-"; s = s.substr(1); }
  const size_t epos = s.find('e');
  const int e10 = std::atoi(s.c_str() + epos + 1);
  std::string digits;
  for (size_t q = 0; q < epos; q++)
    if (s[q] != '.') digits += s[q];
  const int k = (int)digits.size();
  const int n = e10 + 1; /* position of the decimal point */
  if (k <= n && n <= 15) {
    out += digits + std::string(n - k, '0') + ".0";
  } else if (0 < n && n <= 15) {
    out += digits.substr(0, n) + "." + digits.substr(n);
  } else if (-4 < n && n <= 0) {
    out += "0." + std::string(-n, '0') + digits;
  } else {
    out += digits.substr(0, 1);
    if (k > 1) out += "." + digits.substr(1);
    const int ee = n - 1;
    out += ee < 0 ? "e-" : "e+";
    const int ae = ee < 0 ? -ee : ee;
    if (ae < 10) out += "0";
    out += std::to_string(ae);
  }
  return out;
}

}  // namespace ftrh
