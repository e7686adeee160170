// Status-code boundary shared by the C ABI sources: run a body, turn any
// exception into a nonzero status and keep its message for osh_last_error().
#pragma once

#include <exception>
#include <string>

namespace osh {

extern thread_local std::string last_error_message;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (std::exception const& e) {
    last_error_message = e.what();
  } catch (...) {
    last_error_message = "unknown error";
  }
  return 1;
}

}  // namespace osh
