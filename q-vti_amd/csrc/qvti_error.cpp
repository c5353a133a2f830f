// qvti_error.cpp -- the thread-local last-error message of the C ABI (host
// code only, so the host helpers build and run without a GPU toolchain).
#include <cstdarg>
#include <cstdio>
#include <string>

#include "qvti.h"
#include "qvti_internal.h"

static thread_local std::string g_err;

int qvti_fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void qvti_clear_error() { g_err.clear(); }

extern "C" const char* qvti_last_error(void) { return g_err.c_str(); }
