// Error reporting shared by the host sources of the library: sets the
// thread-local message returned by gr_last_error and passes the status on.
#pragma once

#include <string>

#include "gorender_hip.h"

namespace gr_internal {
gr_status fail(gr_status s, const std::string& msg);
}
