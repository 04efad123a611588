// Stand-in for the argument-parser header MISA-MD core ships (args.hpp): just
// the two types the hook file names, so the hook file compiles unchanged.
#pragma once
#include <string>

namespace args {
class ArgumentParser {};
struct Matcher {
  Matcher(char, std::string) {}
};
template <class T>
class ValueFlag {
 public:
  ValueFlag(ArgumentParser &, const std::string &, const std::string &, Matcher) {}
};
}  // namespace args
