// The exception guard every libfot entry point runs its body through: a C++
// exception becomes a status and a message, never an exception across the C
// ABI.  Prints "ok" and exits 0 when every case holds.
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>

#include "fot_internal.hpp"

static int check(bool ok, const char *what) {
  if (!ok) std::printf("FAIL %s (message: %s)\n", what, fot_last_error_message());
  return ok ? 0 : 1;
}

int main() {
  int bad = 0;
  fot_status st = fot::guarded("case_a", []() -> fot_status { throw std::bad_alloc(); });
  bad += check(st == FOT_ERR_OUT_OF_MEMORY &&
                   std::strstr(fot_last_error_message(), "case_a: out of host memory"),
               "bad_alloc");
  st = fot::guarded("case_b", []() -> fot_status { throw std::length_error("too long"); });
  bad += check(st == FOT_ERR_INVALID_ARGUMENT &&
                   std::strstr(fot_last_error_message(), "case_b: too long"),
               "std::exception");
  st = fot::guarded("case_c", []() -> fot_status { throw 42; });
  bad += check(st == FOT_ERR_INVALID_ARGUMENT &&
                   std::strstr(fot_last_error_message(), "case_c: unexpected exception"),
               "other exception");
  st = fot::guarded("case_d", []() -> fot_status { return FOT_ERR_UNSUPPORTED; });
  bad += check(st == FOT_ERR_UNSUPPORTED, "status passes through");
  // an entry point's own argument check still reports as before
  bad += check(fot_loop_create(nullptr, nullptr, nullptr, nullptr) == FOT_ERR_INVALID_ARGUMENT,
               "entry point status");
  if (!bad) std::printf("ok\n");
  return bad ? 1 : 0;
}
