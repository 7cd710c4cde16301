// shkz_error.hpp -- error reporting shared by the library's translation units: records the
// thread's last error message (read by shkz_last_error) and returns the status code.
#pragma once
#include <cstdarg>
#include <cstdio>

__attribute__((visibility("hidden"))) int shkz_set_error(int code, const char *msg);

static inline int shkz_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return shkz_set_error(code, buf);
}
