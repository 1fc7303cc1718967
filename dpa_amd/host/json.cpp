#include "json.hpp"

#include <cctype>
#include <cstdlib>

namespace dpa_hip {

class json_reader {
 public:
  explicit json_reader(const std::string& t) : t_(t) {}

  json document() {
    json v = value();
    ws();
    if (i_ != t_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& t_;
  size_t i_ = 0;

  [[noreturn]] void fail(const std::string& what) {
    throw json_error("json: " + what + " at offset " + std::to_string(i_));
  }
  void ws() {
    while (i_ < t_.size() && std::isspace(static_cast<unsigned char>(t_[i_]))) ++i_;
  }
  bool eat(char c) {
    ws();
    if (i_ < t_.size() && t_[i_] == c) { ++i_; return true; }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail(std::string("expected '") + c + "'");
  }
  bool word(const char* w) {
    const std::string s(w);
    if (t_.compare(i_, s.size(), s) == 0) { i_ += s.size(); return true; }
    return false;
  }

  std::string str() {
    expect('"');
    std::string out;
    while (true) {
      if (i_ >= t_.size()) fail("unterminated string");
      const char c = t_[i_++];
      if (c == '"') break;
      if (c != '\\') { out += c; continue; }
      if (i_ >= t_.size()) fail("bad escape");
      const char e = t_[i_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          if (i_ + 4 > t_.size()) fail("bad \\u escape");
          const unsigned cp = std::stoul(t_.substr(i_, 4), nullptr, 16);
          i_ += 4;
          if (cp < 0x80) out += static_cast<char>(cp);
          else if (cp < 0x800) { out += static_cast<char>(0xC0 | (cp >> 6)); out += static_cast<char>(0x80 | (cp & 0x3F)); }
          else { out += static_cast<char>(0xE0 | (cp >> 12)); out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F)); out += static_cast<char>(0x80 | (cp & 0x3F)); }
          break;
        }
        default: fail("bad escape");
      }
    }
    return out;
  }

  json value() {
    ws();
    if (i_ >= t_.size()) fail("unexpected end");
    json v;
    const char c = t_[i_];
    if (c == '{') {
      ++i_;
      v.kind_ = json::kind::object;
      if (eat('}')) return v;
      do {
        ws();
        std::string k = str();
        expect(':');
        v.object_[k] = value();
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      ++i_;
      v.kind_ = json::kind::array;
      if (eat(']')) return v;
      do {
        v.array_.push_back(value());
      } while (eat(','));
      expect(']');
    } else if (c == '"') {
      v.kind_ = json::kind::string;
      v.string_ = str();
    } else if (word("true")) {
      v.kind_ = json::kind::boolean;
      v.bool_ = true;
    } else if (word("false")) {
      v.kind_ = json::kind::boolean;
      v.bool_ = false;
    } else if (word("null")) {
      v.kind_ = json::kind::null;
    } else {
      const char* b = t_.c_str() + i_;
      char* e = nullptr;
      v.number_ = std::strtod(b, &e);
      if (e == b) fail("bad value");
      i_ += static_cast<size_t>(e - b);
      v.kind_ = json::kind::number;
    }
    return v;
  }
};

json json::parse(const std::string& text) { return json_reader(text).document(); }

const json& json::operator[](const std::string& key) const {
  if (kind_ != kind::object) throw json_error("json: not an object (key '" + key + "')");
  auto it = object_.find(key);
  if (it == object_.end()) throw json_error("json: missing key '" + key + "'");
  return it->second;
}

const json& json::operator[](size_t i) const {
  if (kind_ != kind::array) throw json_error("json: not an array");
  if (i >= array_.size()) throw json_error("json: index out of range");
  return array_[i];
}

size_t json::size() const {
  if (kind_ == kind::array) return array_.size();
  if (kind_ == kind::object) return object_.size();
  return 0;
}

bool json::as_bool() const {
  if (kind_ != kind::boolean) throw json_error("json: not a boolean");
  return bool_;
}

double json::as_number() const {
  if (kind_ != kind::number) throw json_error("json: not a number");
  return number_;
}

const std::string& json::as_string() const {
  if (kind_ != kind::string) throw json_error("json: not a string");
  return string_;
}

}  // namespace dpa_hip
