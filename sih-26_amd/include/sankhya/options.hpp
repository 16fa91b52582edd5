// SPDX-License-Identifier: Apache-2.0
// Option registry: one table of typed, range-checked rows (reference:
// include/sankhya/options.hpp:20-111, src/util/options.cpp). Holds the rows the PDHG path and
// its dispatcher read, plus this engine's own (pdhg_precision, pdhg_chunk).
#pragma once

#include <cstdint>
#include <string>
#include <variant>
#include <vector>

namespace sankhya {

enum class OptionType { Bool, Int, Double, String };
using OptionValue = std::variant<bool, std::int64_t, double, std::string>;

struct OptionSpec {
  std::string name;
  OptionType type = OptionType::String;
  OptionValue default_value;
  std::string description;
  double min_value = 0.0;
  double max_value = 0.0;
  std::vector<std::string> choices;
  std::string planned_for{};
  bool case_sensitive = false;
  [[nodiscard]] bool implemented() const { return planned_for.empty(); }
};

class Options {
 public:
  Options();
  static const std::vector<OptionSpec>& registry();
  static const OptionSpec* find_spec(const std::string& name);
  bool set_from_string(const std::string& name, const std::string& text, std::string* error);
  void set_bool(const std::string& name, bool value);
  void set_int(const std::string& name, std::int64_t value);
  void set_double(const std::string& name, double value);
  void set_string(const std::string& name, const std::string& value);
  [[nodiscard]] bool get_bool(const std::string& name) const;
  [[nodiscard]] std::int64_t get_int(const std::string& name) const;
  [[nodiscard]] double get_double(const std::string& name) const;
  [[nodiscard]] const std::string& get_string(const std::string& name) const;
  [[nodiscard]] static bool exists(const std::string& name);
  [[nodiscard]] bool is_modified(const std::string& name) const;
  [[nodiscard]] std::vector<std::string> modified_names() const;
  [[nodiscard]] std::string value_as_string(const std::string& name) const;

 private:
  [[nodiscard]] std::size_t index_of(const std::string& name) const;
  std::vector<OptionValue> values_;
};

}  // namespace sankhya
