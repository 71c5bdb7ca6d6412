// Error reporting shared by the C ABI translation units.
#pragma once

#include <string>

#include "lpm_bve.h"

namespace lpm {
// Record `msg` as this thread's last error and return `code`.
int fail(int code, const std::string& msg);
// Clear this thread's last error and return LPM_OK.
int ok();
}  // namespace lpm
