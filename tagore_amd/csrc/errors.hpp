// errors.hpp -- the thread-local last-error message behind tg_last_error().
#pragma once
#include <string>

namespace tg {
// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);
}  // namespace tg
