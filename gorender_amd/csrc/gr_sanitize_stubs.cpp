// Host-only build of the library's CPU code (the OBJ/MTL/PNG parsers and the
// host float32 math) under AddressSanitizer and UndefinedBehaviorSanitizer:
// the error-message plumbing, and stand-ins for the two device entry points
// the asset loader calls, which fail with GR_ERR_UNSUPPORTED.
#include <string>

#include "gorender_hip.h"
#include "gr_error.h"

namespace {
thread_local std::string g_last_error;
}

namespace gr_internal {
gr_status fail(gr_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}
}  // namespace gr_internal

extern "C" {

const char* gr_last_error(void) { return g_last_error.c_str(); }

gr_status gr_texture_create(gr_context*, const gr_texture_desc*, const uint8_t*, int32_t*) {
    return gr_internal::fail(GR_ERR_UNSUPPORTED, "host-only sanitizer build: no device");
}

gr_status gr_mesh_create(gr_context*, const gr_mesh_desc*, int32_t*) {
    return gr_internal::fail(GR_ERR_UNSUPPORTED, "host-only sanitizer build: no device");
}

}  // extern "C"
