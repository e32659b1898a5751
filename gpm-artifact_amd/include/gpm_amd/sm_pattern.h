// Pattern names, as the reference's Pattern class (src/utils/sm_pattern.h:
// 22-38): solvers dispatch on "P1".."P8" and "Motif4".
#pragma once

#include <string>

namespace project_AntiRF {

class Pattern {
 public:
  Pattern() : Pattern("") {}
  Pattern(std::string name) : name_(std::move(name)) {}
  bool is_rectangle() const { return name_ == "rectangle"; }
  bool is_pentagon() const { return name_ == "pentagon"; }
  bool is_house() const { return name_ == "house"; }
  std::string get_name() const { return name_; }
  // gpm_pattern_id of the name, 0 when unknown.
  int id() const {
    static const char* names[] = {"P1", "P2", "P3", "P4", "P5", "P6", "P7", "P8", "Motif4"};
    for (int i = 0; i < 9; ++i)
      if (name_ == names[i]) return i + 1;
    return 0;
  }

 private:
  std::string name_;
};

}  // namespace project_AntiRF
