// error.cpp -- the reference's ErrorStatus / traceback convention
// (src/error.h:14-88, src/error.c:25-492): an error carries a heap string
// that every WRAP_TRACEBACK level prepends a "File ..., line ... in ..."
// frame to; the caller prints it with print_and_free_traceback or frees it
// with free_traceback.  Messages use the same layout and error-type names.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "gs_internal.h"

namespace {

const char *kReset = "\033[0m";
const char *kCyan = "\033[0;36m";
const char *kPurple = "\033[0;35m";
const char *kPurpleBold = "\033[1;95m";

const char *type_name(int code, int *normalised)
{
    *normalised = code;
    switch (code) {
        case GRAV_FAILURE: return "Failure";
        case GRAV_VALUE_ERROR: return "ValueError";
        case GRAV_POINTER_ERROR: return "PointerError";
        case GRAV_MEMORY_ERROR: return "MemoryError";
        case GRAV_OS_ERROR: return "OSError";
        case GRAV_NOT_IMPLEMENTED_ERROR: return "NotImplementedError";
        default: *normalised = GRAV_UNKNOWN_ERROR; return "UnknownError";
    }
}

std::string vformat(const char *fmt, va_list ap)
{
    va_list ap2;
    va_copy(ap2, ap);
    int n = vsnprintf(nullptr, 0, fmt, ap2);
    va_end(ap2);
    if (n < 0) return std::string();
    std::string s(static_cast<size_t>(n) + 1, '\0');
    vsnprintf(&s[0], s.size(), fmt, ap);
    s.resize(static_cast<size_t>(n));
    return s;
}

char *heap_copy(const std::string &s)
{
    char *p = static_cast<char *>(malloc(s.size() + 1));
    if (p) memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

ErrorStatus make_error(const char *file, int line, const char *func, int code,
                       const std::string &msg)
{
    ErrorStatus st;
    const char *type = type_name(code, &st.return_code);
    std::string tb = std::string("    File ") + kCyan + "\"" + file + "\"" + kReset + ", line " +
                     kCyan + std::to_string(line) + kReset + " in " + kCyan + func + kReset +
                     "\n" + kPurpleBold + type + kReset + ": " + kPurple + msg + kReset + "\n";
    st.traceback = heap_copy(tb);
    st.traceback_code_ = st.traceback ? GRAV_TRACEBACK_SUCCESS : GRAV_TRACEBACK_MALLOC_FAILED;
    return st;
}

}  // namespace

extern "C" {

ErrorStatus make_success_error_status(void)
{
    ErrorStatus st;
    st.return_code = GRAV_SUCCESS;
    st.traceback = nullptr;
    st.traceback_code_ = GRAV_TRACEBACK_NOT_INITIALIZED;
    return st;
}

void raise_warning(const char *file, const int line, const char *func, const char *msg)
{
    fprintf(stderr, "\033[1;33mWarning:%s In %s\"%s\"%s, line %s%d%s in %s%s%s:\n    %s%s%s\n",
            kReset, kCyan, file, kReset, kCyan, line, kReset, kCyan, func, kReset, kPurple, msg,
            kReset);
}

ErrorStatus raise_warning_fmt(const char *file, const int line, const char *func,
                              const char *format, ...)
{
    va_list ap;
    va_start(ap, format);
    std::string msg = vformat(format, ap);
    va_end(ap);
    raise_warning(file, line, func, msg.c_str());
    return make_success_error_status();
}

ErrorStatus raise_error(const char *file, const int line, const char *func, const int code,
                        const char *msg)
{
    return make_error(file, line, func, code, msg ? msg : "");
}

ErrorStatus raise_error_fmt(const char *file, const int line, const char *func, const int code,
                            const char *format, ...)
{
    va_list ap;
    va_start(ap, format);
    std::string msg = vformat(format, ap);
    va_end(ap);
    return make_error(file, line, func, code, msg);
}

ErrorStatus traceback(ErrorStatus st, const char *call, const char *file, const int line,
                      const char *func)
{
    if (st.return_code == GRAV_SUCCESS || st.traceback_code_ != GRAV_TRACEBACK_SUCCESS)
        return st;
    std::string frame = std::string("    File ") + kCyan + "\"" + file + "\"" + kReset +
                        ", line " + kCyan + std::to_string(line) + kReset + " in " + kCyan +
                        func + kReset + "\n        " + kPurple + call + kReset + "\n" +
                        (st.traceback ? st.traceback : "");
    char *p = heap_copy(frame);
    if (!p) {
        free(st.traceback);
        st.traceback = nullptr;
        st.traceback_code_ = GRAV_TRACEBACK_MALLOC_FAILED;
        return st;
    }
    free(st.traceback);
    st.traceback = p;
    return st;
}

void free_traceback(ErrorStatus *st)
{
    if (st->traceback) {
        free(st->traceback);
        st->traceback = nullptr;
    }
}

void print_and_free_traceback(ErrorStatus *st)
{
    fprintf(stderr, "\n\033[1;31mTraceback%s \033[5;31m(most recent call last):%s\n", kReset,
            kReset);
    switch (st->traceback_code_) {
        case GRAV_TRACEBACK_NOT_INITIALIZED:
            fputs("    Something went wrong. Traceback not initialized.\n", stderr);
            break;
        case GRAV_TRACEBACK_SUCCESS:
        case GRAV_TRACEBACK_TRUNCATED:
            if (st->traceback) fputs(st->traceback, stderr);
            if (st->traceback_code_ == GRAV_TRACEBACK_TRUNCATED)
                fputs("\n    Something went wrong. Traceback are truncated.\n", stderr);
            free(st->traceback);
            st->traceback = nullptr;
            break;
        case GRAV_TRACEBACK_MALLOC_FAILED:
            fputs("    Something went wrong. Failed to allocate memory for traceback.\n", stderr);
            break;
        case GRAV_TRACEBACK_SNPRINTF_FAILED:
            fputs("    Something went wrong. Failed to write to traceback.\n", stderr);
            break;
    }
}

}  // extern "C"

namespace gs {

ErrorStatus hip_error(hipError_t e, const char *what, const char *file, int line,
                      const char *func)
{
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                         ? GRAV_MEMORY_ERROR
                         : GRAV_FAILURE;
    return raise_error_fmt(file, line, func, code, "HIP call failed: %s -> %s (%d)", what,
                           hipGetErrorString(e), static_cast<int>(e));
}

}  // namespace gs
