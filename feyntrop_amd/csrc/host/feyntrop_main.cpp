/*
 * feyntrop_main.cpp -- the `feyntrop` executable: reads the low-level JSON
 * problem from stdin, writes the result JSON line to stdout and log lines to
 * stderr (srcs/feyntrop.cpp:20-281).  Exit status 0 on success, 1 on any
 * error (the reference aborts; callers check only for nonzero).
 * FEYNTROP_AMD_DEVICE selects the HIP device (default 0).
 */
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <iterator>
#include <string>

#include "../../../include/feyntrop_amd.h"

int main() {
  std::string in((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
  int dev = 0;
  if (const char* d = getenv("FEYNTROP_AMD_DEVICE")) dev = atoi(d);
  /* Without a device the input is still parsed and validated, so input
   * errors are reported the same way on any host. */
  ftr_ctx* ctx = nullptr;
  if (ftr_create(dev, &ctx) != FTR_OK) ctx = nullptr;
  char *out = nullptr, *log = nullptr;
  const int rc = ftr_run_json(ctx, in.c_str(), &out, &log);
  if (log) std::cerr << log << std::flush;
  if (rc == FTR_OK && out) std::cout << out << std::endl;
  ftr_free_string(out);
  ftr_free_string(log);
  ftr_destroy(ctx);
  return rc == FTR_OK ? 0 : 1;
}
