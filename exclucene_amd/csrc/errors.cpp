// errors.cpp -- the calling thread's last error message (exc_last_error) and
// the helper every entry point returns its status through.  Host-only, so the
// host builder can be built without the HIP runtime (the sanitizer build).
#include <cstdarg>
#include <cstdio>

#include "../../include/exclucene_amd.h"

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* exc_last_error(void) { return g_err; }
