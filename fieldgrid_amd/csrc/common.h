// Shared host helpers of libfieldgrid: the thread-local error text behind fg_last_error().
#pragma once

#include <string>

namespace fg {

// Records `msg` as this thread's last error and returns `status` (the reference returns
// Result<_, String>; the C-ABI returns a status and keeps the String here).
int set_error(int status, const std::string &msg);

}  // namespace fg
