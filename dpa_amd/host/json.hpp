// json.hpp -- the small JSON reader the configuration parser needs
// (objects, arrays, strings, numbers, booleans, null).
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace dpa_hip {

class json {
 public:
  enum class kind { null, boolean, number, string, array, object };

  json() = default;
  static json parse(const std::string& text);

  kind type() const { return kind_; }
  bool is_null() const { return kind_ == kind::null; }
  bool contains(const std::string& key) const {
    return kind_ == kind::object && object_.count(key) != 0;
  }
  const json& operator[](const std::string& key) const;
  const json& operator[](size_t i) const;
  size_t size() const;

  bool as_bool() const;
  double as_number() const;
  const std::string& as_string() const;

 private:
  kind kind_ = kind::null;
  bool bool_ = false;
  double number_ = 0.0;
  std::string string_;
  std::vector<json> array_;
  std::map<std::string, json> object_;

  friend class json_reader;
};

struct json_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

}  // namespace dpa_hip
