// dgpu_error.hpp — the library's last-error slot, shared by its translation units
// (dgpu_last_error returns the message of the last failing call on this thread).
#pragma once
#include <string>

namespace dgpu {
int set_error(int code, const std::string& msg);
}
