// Error handling shared by the C ABI translation units: every entry point runs
// its body under guarded(), which turns exceptions into a nonzero status and a
// per-thread message instead of letting them cross the C boundary.
#pragma once
#include <exception>
#include <stdexcept>
#include <string>

namespace dkt {
extern thread_local std::string g_last_error;

template <typename F>
int guarded(F &&f) {
  try {
    g_last_error.clear();
    f();
    return 0;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return 1;
  } catch (...) {
    g_last_error = "unknown error";
    return 1;
  }
}
}  // namespace dkt
