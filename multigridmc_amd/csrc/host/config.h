// Reader for the libconfig file format used by the reference's parameter and
// measurement files (parameters_template.cfg, measurements_template.cfg):
// groups { }, arrays [ ], lists ( ), integers (decimal / hex, optional L),
// floats, booleans, strings (adjacent literals concatenated), "=" or ":"
// separators, ";" or "," terminators, and #, //, /* */ comments.
// Lookups mirror libconfig::Setting::lookup (dotted paths); a missing
// setting or a type mismatch throws ConfigError, as libconfig's
// SettingNotFoundException / SettingTypeException do.
#ifndef MGMC_HOST_CONFIG_H
#define MGMC_HOST_CONFIG_H

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mgmc_host {

struct ConfigError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Setting {
  enum Kind { Int, Float, Bool, String, Array, List, Group };
  Kind kind = Group;
  std::string name;
  int64_t i = 0;
  double d = 0.0;
  bool b = false;
  std::string s;
  std::vector<Setting> items;  // Array / List elements, Group members (named)

  const Setting& lookup(const std::string& path) const;  // dotted path below this group
  bool exists(const std::string& path) const;
  const Setting& operator[](size_t idx) const;           // Array / List element
  size_t length() const { return items.size(); }

  int64_t as_int() const;
  double as_double() const;  // integers convert, as with libconfig's auto-conversion
  bool as_bool() const;
  const std::string& as_string() const;
  std::vector<double> as_doubles() const;  // Array / List of numbers
};

// Parse a file / a string; the root is a Group.
Setting read_config_file(const std::string& filename);
Setting parse_config(const std::string& text, const std::string& origin = "<string>");

}  // namespace mgmc_host

#endif
