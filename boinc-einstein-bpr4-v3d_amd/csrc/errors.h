// errors.h -- thread-local error message behind brp_last_error().
#pragma once
#include <string>

namespace brp {
int fail(int code, const std::string& message);
}
