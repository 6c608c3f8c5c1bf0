// gcs_error.hpp -- the library's per-thread last-error message, shared by the
// engine and the job server (read back through gcs_last_error()).
#pragma once

namespace gcs_detail {
// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const char* msg);
}  // namespace gcs_detail
