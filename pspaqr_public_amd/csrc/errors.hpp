// Message of the last failed C-ABI call on this thread (spaqr_last_error with no tree).
#pragma once
#include <string>

namespace spaqr {
inline std::string& api_error() {
    static thread_local std::string e;
    return e;
}
}  // namespace spaqr
