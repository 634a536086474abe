// Shared host-side helpers: thread-local error message, HIP status checks.
#pragma once

#include <string>

namespace sbi {

void set_error(const std::string& msg);

}  // namespace sbi
