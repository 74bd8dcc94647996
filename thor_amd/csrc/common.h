// Shared host-side helpers of the library.
#pragma once
#include <string>

namespace thor {
void set_error(const std::string &msg);
inline int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}
} // namespace thor
