/* Thread-local last-error message of the C ABI (mgx_last_error). */
#pragma once

#include <cstdarg>
#include <cstdio>

namespace mgx {

inline char *error_buffer()
{
    static thread_local char buf[512] = {0};
    return buf;
}

inline void clear_error() { error_buffer()[0] = 0; }

inline int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(error_buffer(), 512, fmt, ap);
    va_end(ap);
    return code;
}

} // namespace mgx
