// Host driver for tests/test_math.py: evaluates the kernels' elementary
// functions (feyntrop_amd/csrc/kernels/ftr_math.hpp, compiled as plain C++)
// on arguments read from stdin, one per line: "<fn> <hex double> [<hex double>]",
// and prints the results as hex doubles.
#include <cstdio>
#include <cstring>

#include "ftr_math.hpp"

int main() {
  char fn[32];
  double x, y;
  while (std::scanf("%31s %la", fn, &x) == 2) {
    if (!std::strcmp(fn, "sincos")) {
      double s, c;
      ftrm::sincos_fast(x, &s, &c);
      std::printf("%a %a\n", s, c);
    } else if (!std::strcmp(fn, "atan2")) {
      if (std::scanf("%la", &y) != 1) return 1;
      std::printf("%a\n", ftrm::atan2_fast(x, y));
    } else if (!std::strcmp(fn, "log_pos")) {
      std::printf("%a\n", ftrm::log_pos(x));
    } else if (!std::strcmp(fn, "log_any")) {
      std::printf("%a\n", ftrm::log_any(x));
    } else if (!std::strcmp(fn, "rcp")) {
      std::printf("%a\n", ftrm::rcp_nr(x));
    } else {
      return 2;
    }
  }
  return 0;
}
