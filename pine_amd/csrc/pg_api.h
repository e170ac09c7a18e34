// Internal helpers shared by the library's translation units.
#pragma once

#include "../../include/pinehip.h"
#include "pg_vec.h"

#include <string>

namespace pg {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

// Changes whenever the accel's hierarchy or launch configuration changes
// (work recorded in a HIP graph against the old one must be recorded again).
int64_t accel_generation(const pg_accel* a);

}  // namespace pg

inline int pg_fail(int code, const std::string& msg) { return pg::fail(code, msg); }
