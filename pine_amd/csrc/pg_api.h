// Internal helpers shared by the library's translation units.
#pragma once

#include "../../include/pinehip.h"
#include "pg_vec.h"

#include <string>

namespace pg {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

}  // namespace pg

inline int pg_fail(int code, const std::string& msg) { return pg::fail(code, msg); }
