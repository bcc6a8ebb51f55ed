// vx_common.hpp -- error plumbing shared by the host-side translation units of
// libvixen_hip.so. Nothing throws across the C ABI: each entry point records a
// message in a thread-local buffer and returns a vx_status.
#pragma once

#include "vixen_hip.h"

#include <string>

namespace vx {

// Records msg as this thread's last error and returns code.
vx_status fail(vx_status code, const std::string& msg);

// Clears this thread's last error.
void clear_error();

} // namespace vx
