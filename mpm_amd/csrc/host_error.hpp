// host_error.hpp -- the error message of C-ABI calls that have no handle to
// keep it on (mpm_last_error(NULL)); one per calling thread.
#pragma once

namespace mpm_host {
void set_global_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* global_error();
}  // namespace mpm_host
