// The message behind wpt_last_error(), set by every failing entry point of
// libwpt.so; returns `code`.
#pragma once

#include <string>

namespace wpt {
int setLastError(int code, const std::string& message);
}
